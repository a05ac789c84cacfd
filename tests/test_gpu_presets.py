"""Device presets (plife_set_positions / plife_set_types) against the
oracle's restatement (oracle/presets.py, pinned in test_presets.py).

Positions: the device's sin/cos/log/sqrt may differ from numpy's by an ulp,
so fixed-point positions must agree within 2 units of 2^-32 (periodic
distance); types, velocities (0) and everything else exactly."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import presets as op

pytestmark = pytest.mark.gpu

T = 6


@pytest.fixture(scope="module")
def eng():
    from particle_life import Engine

    e = Engine(0)
    yield e
    e.close()


def fixed_state(st):
    p = np.round(st["pos"] * 2.0 ** 32) % 2.0 ** 32
    return p.astype(np.uint64).astype(np.uint32)


def close_fixed(a, b, wrap):
    d = (a.astype(np.int64) - b.astype(np.int64))
    if wrap:
        d = (d + 2**31) % 2**32 - 2**31
    return np.abs(d).max() if d.size else 0


@pytest.mark.parametrize("dim,wrap", [(2, True), (2, False), (3, True)])
def test_position_presets(eng, dim, wrap):
    from particle_life import worlds

    w = worlds.uniform_world(301, 5_000, T, rmax=0.03, dt=0.01, dim=dim, wrap=wrap)
    for i, name in enumerate(op.POSITION_PRESETS):
        print(f"\n[presets] position {name} dim={dim} wrap={wrap}", flush=True)
        eng.load_world(w)
        eng.step(3)                       # slots no longer in upload order; velocities nonzero
        before = eng.state()
        seed = 1000 + i
        eng.set_positions(name, seed)
        st = eng.state()
        px, py, pz = op.set_positions(i, before["types"], T, seed, dim, wrap)
        got = fixed_state(st)
        assert close_fixed(got[:, 0], px, wrap) <= 2, name
        assert close_fixed(got[:, 1], py, wrap) <= 2, name
        if dim == 3:
            np.testing.assert_array_equal(got[:, 2], pz)
        np.testing.assert_array_equal(st["vel"], 0.0)
        np.testing.assert_array_equal(st["types"], before["types"])
        np.testing.assert_array_equal(st["ids"], before["ids"])
        eng.step(2)                       # the state steps normally afterwards


@pytest.mark.parametrize("dim", [2, 3])
def test_type_presets(eng, dim):
    from particle_life import worlds

    w = worlds.uniform_world(302, 5_000, T, rmax=0.03, dt=0.01, dim=dim)
    for i, name in enumerate(op.TYPE_PRESETS):
        print(f"\n[presets] type {name} dim={dim}", flush=True)
        eng.load_world(w)
        eng.step(4)
        before = eng.state()
        fx = fixed_state(before)
        seed = 2000 + i
        eng.set_types(name, seed)
        st = eng.state()
        want = op.set_types(i, fx[:, 0], fx[:, 1], fx[:, 2], before["vel"].astype(np.float32),
                            before["types"], T, seed, dim)
        np.testing.assert_array_equal(st["types"], want, err_msg=name)
        np.testing.assert_array_equal(st["pos"], before["pos"])
        np.testing.assert_array_equal(st["vel"], before["vel"])


def test_batch_world_presets_leave_other_worlds_alone():
    from particle_life import Batch, worlds

    print("\n[presets] batch", flush=True)
    ws = [worlds.uniform_world(310 + k, 3_000 + 500 * k, T, rmax=0.03, dt=0.01) for k in range(3)]
    b = Batch(0)
    try:
        ids = [b.add_world(w) for w in ws]
        b.step(2)
        b.sync()
        before = [b.state(i) for i in ids]
        b.set_positions(ids[1], "spiral", 7)
        b.set_types(ids[1], "rotate", 8)
        after = [b.state(i) for i in ids]
        for k in (0, 2):
            for key in ("pos", "vel", "types", "ids"):
                np.testing.assert_array_equal(after[k][key], before[k][key])
        px, py, _ = op.set_positions(op.POSITION_PRESETS.index("spiral"), before[1]["types"], T, 7,
                                     2, True)
        got = fixed_state(after[1])
        assert close_fixed(got[:, 0], px, True) <= 2 and close_fixed(got[:, 1], py, True) <= 2
        np.testing.assert_array_equal(after[1]["types"], (before[1]["types"] + 1) % T)
        b.step(2)
        b.sync()
    finally:
        b.close()


def test_gpu_physics_uses_the_device_presets():
    from particle_life import GpuPhysics, ParticleLifeAccelerator, presets

    print("\n[presets] GpuPhysics", flush=True)
    phys = GpuPhysics(ParticleLifeAccelerator(0.3), particleCount=4_000,
                      positionSetter=presets.PresetPositionSetter("ring", seed=5),
                      typeSetter=presets.PresetTypeSetter("slices"))
    try:
        phys.update(2)
        phys.setPositions()
        p = phys.particles
        r = np.hypot(p.position[:, 0] - 0.5, p.position[:, 1] - 0.5)
        assert abs(r.mean() - 0.35) < 0.003 and np.all(p.velocity == 0)
        phys.setTypes()
        p = phys.particles
        n = phys.settings.matrix.size()
        np.testing.assert_array_equal(p.type, np.clip(np.floor(p.position[:, 0] * n), 0, n - 1))
        phys.update(2)
    finally:
        phys.kill()
