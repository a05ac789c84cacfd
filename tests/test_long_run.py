"""Pins the long-run comparison (tests/long_run.py) on the CPU: the type-pair
radial distribution statistic stays within its tolerance for runs that differ
only by tiny perturbations, and fails by far for a run with a planted mistake
in the oracle (matrix transposed, the acted-on particle's row swapped for the
other's) -- so the GPU test built on it can tell a wrong force law."""
from __future__ import annotations

import numpy as np

import oracle
from long_run import chi2_tolerance, error_tolerance, oracle_run, rdf, spread
from parity_util import oracle_settings
from rdf_util import two_sample_chi2


def _world():
    from particle_life import worlds

    return worlds.uniform_world(worlds.config_seed(1), 3000, 6, rmax=0.02, dt=0.02)


def test_rdf_tolerance_accepts_nudged_runs_and_rejects_a_planted_mistake():
    w = _world()
    steps = 60
    ref, h, chi, err = spread(w, steps, nruns=2)
    # a run rounded like the device (FP32 velocities, 2^-32 positions) stays inside
    S = oracle_settings(w)
    pos, vel, typ, ids = w.pos, w.vel, w.types, None
    for _ in range(steps):
        r = oracle.step(S, pos, vel, typ, ids=ids, nthreads=4)
        pos, typ, ids = r["pos"], r["types"], r["ids"]
        vel = r["vel"].astype(np.float32).astype(np.float64)
        pos = np.round(pos * 2.0 ** 32) * 2.0 ** -32
        pos = pos - np.floor(pos)
    dev_like = (pos, typ, ids)
    assert two_sample_chi2(h, rdf(w, pos, typ))[0] <= chi2_tolerance(chi)
    from long_run import max_error
    assert max_error(w, ref, dev_like) <= max(error_tolerance(err), 1e-9)

    # planted mistake: the transposed matrix
    bad = oracle_settings(w)
    bad.matrix = np.ascontiguousarray(w.matrix.T)
    wrong = oracle_run(w, w.pos, steps, settings=bad)
    stat = two_sample_chi2(h, rdf(w, wrong[0], wrong[1]))[0]
    assert stat > 20 * chi2_tolerance(chi), (stat, chi)
