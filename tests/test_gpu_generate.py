"""Device-side seeded world generation equals the host generator
(worlds.uniform_world) bit for bit, and steps identically."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,wrap", [(2, True), (3, True), (2, False)])
def test_generated_equals_host_world(dim, wrap):
    from particle_life import Engine, worlds

    print(f"\n[generate] case=engine dim={dim} wrap={wrap}", flush=True)
    seed, n, t = worlds.config_seed(9, dim), 30_000, 7
    w = worlds.uniform_world(seed, n, t, rmax=0.03, dt=0.01, dim=dim, wrap=wrap)
    a, b = Engine(0), Engine(0)
    a.configure_world(w)
    a.generate(n, seed)
    b.load_world(w)
    sa, sb = a.state(), b.state()
    np.testing.assert_array_equal(sa["pos"], w.pos)              # the host generator's values
    np.testing.assert_array_equal(sa["types"], w.types)
    assert (sa["vel"] == 0).all()
    np.testing.assert_array_equal(sa["ids"], np.arange(n))
    for k in ("pos", "vel", "types", "ids"):
        np.testing.assert_array_equal(sa[k], sb[k])
    a.step(3)
    b.step(3)
    sa, sb = a.state(), b.state()
    for k in ("pos", "vel", "types", "ids"):
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
    a.close()
    b.close()


def test_batch_generated_world():
    from particle_life import Batch, worlds

    print("\n[generate] case=batch", flush=True)
    specs = worlds.config4_specs(6)
    b1, b2 = Batch(0), Batch(0)
    ids1, ids2 = [], []
    for p in specs:
        w = worlds.uniform_world(p["seed"], p["n"], p["ntypes"], rmax=p["rmax"], dt=0.01, beta=p["beta"])
        ids1.append(b1.add_world(w))
        ids2.append(b2.add_generated(n=p["n"], seed=p["seed"], matrix=w.matrix, rmax=w.rmax,
                                     friction=w.friction, dt=w.dt, beta=w.beta))
    b1.step(4)
    b2.step(4)
    for i, j in zip(ids1, ids2):
        s1, s2 = b1.state(i), b2.state(j)
        for k in ("pos", "vel", "types", "ids"):
            np.testing.assert_array_equal(s1[k], s2[k], err_msg=k)
    b1.close()
    b2.close()
