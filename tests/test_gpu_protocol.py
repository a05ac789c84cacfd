"""The rest of the parity protocol on the device: a 100-step C1 run compared
with the oracle's by type-pair radial distribution and position error, each
within a tolerance derived from the oracle's own spread under 2^-32 nudges of
its input; and the reference's stop request (Physics.forceUpdateStop,
backend/Physics.java:118-133,149-151) sent from another thread during a long
update."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

from long_run import chi2_tolerance, error_tolerance, max_error, rdf, spread
from rdf_util import two_sample_chi2

pytestmark = pytest.mark.gpu


def test_c1_hundred_steps_statistically():
    from particle_life import Engine, worlds

    print("\n[protocol] case=c1_100_steps_rdf", flush=True)
    w = worlds.config1()
    ref, h, chi, err = spread(w, 100, nruns=3, nthreads=16)
    with Engine(0) as e:
        e.load_world(w)
        e.step(100)
        g = e.state()
    stat = two_sample_chi2(h, rdf(w, g["pos"], g["types"]))[0]
    dx = max_error(w, ref, (g["pos"], g["types"], g["ids"]))
    print(f"  rdf chi2 {stat:.4g} (tolerance {chi2_tolerance(chi):.4g}); max position error "
          f"{dx:.3e} (tolerance {error_tolerance(err):.3e})", flush=True)
    assert stat <= chi2_tolerance(chi)
    assert dx <= error_tolerance(err)


def _stop_world():
    from particle_life import worlds

    # ~60 interacting neighbours per particle: about 0.1 ms per step
    return worlds.uniform_world(worlds.config_seed(9), 200_000, 6, rmax=0.01, dt=0.01)


def test_stop_from_another_thread_ends_at_a_step_boundary():
    from particle_life import Engine
    from particle_life._native import PLIFE_STOPPED

    print("\n[protocol] case=stop_from_another_thread", flush=True)
    w = _stop_world()
    e = Engine(0)
    e.load_world(w)
    e.step(4)
    e.sync()
    s0 = e.steps_taken()
    out = {}

    def run():
        t = time.perf_counter()
        out["rc"] = e.step(1_000_000)             # minutes of work if never stopped
        out["t"] = time.perf_counter() - t

    th = threading.Thread(target=run)
    th.start()
    time.sleep(0.5)
    t_stop = time.perf_counter()
    e.stop()
    th.join(timeout=120)
    assert not th.is_alive(), "plife_update did not return after the stop"
    returned_after = time.perf_counter() - t_stop
    e.sync()
    k = e.steps_taken() - s0
    print(f"  rc={out.get('rc')} steps before the stop took effect: {k}, update returned "
          f"{returned_after * 1e3:.1f} ms after the request", flush=True)
    assert out["rc"] == PLIFE_STOPPED
    assert 0 < k < 1_000_000
    assert returned_after < 5.0
    stopped = e.state()
    e.close()
    # the state is the whole-step state after 4 + k steps
    with Engine(0) as a:
        a.load_world(w)
        a.step(4 + k)
        alone = a.state()
    for key in ("pos", "vel", "ids", "types"):
        np.testing.assert_array_equal(stopped[key], alone[key], err_msg=key)


def test_stop_before_update_is_cleared():
    """As Physics.updateParticles sets its run flag first (Physics.java:118),
    a stop request sent before an update does not cancel it."""
    from particle_life import Engine

    print("\n[protocol] case=stop_before_update", flush=True)
    with Engine(0) as e:
        e.load_world(_stop_world())
        e.stop()
        s0 = e.steps_taken()
        assert e.step(6) == 0
        e.sync()
        assert e.steps_taken() - s0 == 6
