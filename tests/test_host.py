"""Host-side logic that needs no GPU: the Physics API mirror's value types,
the seeded world generators and the accelerator as the app defines it."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from particle_life import physics, worlds


def test_splitmix64_known_answers():
    # SplitMix64 from state 0: the first three outputs (Vigna's reference generator)
    g = worlds.GAMMA
    with np.errstate(over="ignore"):        # k * gamma wraps modulo 2^64 by design
        z = worlds._mix(np.array([g, 2 * g, 3 * g], dtype=np.uint64))
    assert [hex(int(v)) for v in z] == ["0xe220a8397b1dcdaf", "0x6e789e6aa1b965f4",
                                        "0x6c45d188009454f"]


def test_worlds_are_deterministic_and_on_the_grid():
    a = worlds.config1()
    b = worlds.config1()
    np.testing.assert_array_equal(a.pos, b.pos)
    np.testing.assert_array_equal(a.types, b.types)
    assert np.all((a.pos[:, :2] >= 0) & (a.pos[:, :2] < 1)) and np.all(a.pos[:, 2] == 0)
    assert np.all(a.pos * 2 ** 24 == np.round(a.pos * 2 ** 24))
    assert np.all(a.matrix * 2 ** 23 == np.round(a.matrix * 2 ** 23))
    assert a.types.min() == 0 and a.types.max() == 5
    assert abs(a.matrix.mean()) < 0.3
    c3 = worlds.gaussian_blob_world(3, 20_000, 12, rmax=0.01, dt=0.01, beta=0.25)
    assert np.all((c3.pos >= 0) & (c3.pos < 1))
    # clustered: far more particles per occupied cell than uniform
    keys = (c3.pos * 100).astype(int) @ np.array([1, 100, 10000])
    assert np.bincount(keys).max() >= 8    # uniform would give ~0.02 per cell


def test_config_mapping():
    w = worlds.config2(1000)
    assert w.rmax == 0.005 and w.dt == 0.01 and w.ntypes == 8 and w.beta == 0.3
    assert oracle.cells_per_axis(w.rmax) == 200
    assert oracle.friction_factor(w.friction, 0.04) == pytest.approx(0.5)


def test_accelerator_mirror_matches_oracle():
    acc = physics.ParticleLifeAccelerator(0.3)
    rng = np.random.default_rng(5)
    for _ in range(200):
        rel = rng.uniform(-0.7, 0.7, 3)
        rel[2] = 0
        a = rng.uniform(-1, 1)
        np.testing.assert_allclose(acc.accelerate(a, rel.copy()), oracle.accelerate(a, 0.3, rel),
                                   rtol=1e-15, atol=1e-17)


def test_matrix_and_settings_value_semantics():
    m = physics.DefaultMatrix(3)
    m.set(0, 2, 0.5)
    c = m.deepCopy()
    assert c == m and c.get(0, 2) == 0.5
    c.set(0, 2, -1.0)
    assert m.get(0, 2) == 0.5 and c != m
    s = physics.PhysicsSettings(matrix=m)
    t = s.deepCopy()
    assert t == s
    t.rmax = 0.03
    assert t != s
    # the default dt is the float literal 0.02f widened to double
    assert physics.PhysicsSettings().dt == float(np.float32(0.02)) != 0.02


def test_gpu_physics_rejects_other_accelerators():
    class Custom(physics.Accelerator):
        def accelerate(self, a, pos):
            return pos

    with pytest.raises(NotImplementedError):
        physics.GpuPhysics(Custom())


def test_served_trace_generator():
    from particle_life import worlds
    from particle_life.serve import footprint

    js = worlds.config5_jobs(300, rate=40.0)
    arr = np.array([j["arrival"] for j in js])
    assert arr[0] == 0.0 and np.all(np.diff(arr) >= 0)
    assert 3.0 < arr[-1] < 12.0                                  # ~300 / 40 s
    n = np.array([j["n"] for j in js])
    assert n.min() >= 2000 and n.max() <= 262144
    assert 35_000 < n.mean() < 70_000                            # log-uniform mean ~53k
    assert {j["dim"] for j in js} == {2, 3}
    assert all(100 <= j["steps"] <= 2000 for j in js)
    assert all(4 <= j["ntypes"] <= 10 and 0.01 <= j["rmax"] <= 0.03 for j in js)
    w = worlds.job_world(js[0])
    assert w.n == js[0]["n"] and w.dim == js[0]["dim"]
    assert footprint(1000, 0.02, 2) == 1000 * 90 + 12 * 2500


def test_batched_workload_generator():
    from particle_life import worlds

    specs = worlds.config4_specs(512)
    assert specs[:10] == worlds.config4_specs(10)                # prefix-stable
    n = np.array([p["n"] for p in specs])
    assert n.min() >= 1000 and n.max() <= 65536
    assert 7.0e6 < n.sum() < 8.8e6                               # ~7.9M expected
    assert {p["ntypes"] for p in specs} == set(range(3, 13))
    assert all(0.01 <= p["rmax"] <= 0.04 and 0.2 <= p["beta"] <= 0.4 for p in specs)
