"""The drop-in mirror of the reference's Physics API on the device: update()
against the oracle, the reference's setters (setParticleCount, setMatrixSize,
ensureTypes, setPositions), snapshots and save-state round trips."""
from __future__ import annotations

import io

import numpy as np
import pytest

import oracle
from parity_util import VEL_ABS, VEL_REL

pytestmark = pytest.mark.gpu


def _physics(n=5000, seed=5):
    from particle_life import (DefaultMatrixGenerator, DefaultPositionSetter, DefaultTypeSetter,
                               GpuPhysics, ParticleLifeAccelerator)

    rng = np.random.default_rng(seed)
    return GpuPhysics(ParticleLifeAccelerator(0.3), DefaultPositionSetter(rng),
                      DefaultMatrixGenerator(rng), DefaultTypeSetter(rng), particleCount=n)


def test_update_matches_oracle():
    print("\n[physics] case=update_vs_oracle", flush=True)
    ph = _physics()
    ph.settings.dt = 0.02                      # the app's per-frame dt (app/Main.java:292-295)
    ph.update(0)                               # upload: the device rounds positions to 2^-32
    p0 = ph.particles
    s = ph.settings
    S = oracle.Settings(matrix=s.matrix.to_array(), rmax=s.rmax, friction=s.friction, dt=s.dt,
                        beta=ph.accelerator.beta, force=s.force, wrap=s.wrap, dim=2)
    ref = oracle.step(S, p0.position, p0.velocity, p0.type, nthreads=8)
    for _ in range(1):
        ph.update()
    p1 = ph.particles
    np.testing.assert_array_equal(p1.ids, ref["ids"], err_msg="particle order")
    np.testing.assert_array_equal(p1.type, ref["types"])
    dv = np.abs(p1.velocity - ref["vel"])
    assert not (dv > VEL_ABS + VEL_REL * np.abs(ref["vel"])).any()
    d = np.abs(p1.position - ref["pos"])
    d = np.minimum(d, 1.0 - d)
    assert d.max() < 2.0 ** -32 + s.dt * (VEL_ABS + VEL_REL * np.abs(ref["vel"]).max())
    ph.kill()


def test_reference_setters_and_snapshot():
    print("\n[physics] case=setters", flush=True)
    ph = _physics(3000)
    ph.update(2)
    ph.setParticleCount(4500)
    assert len(ph.particles) == 4500
    ph.update(2)
    ph.setMatrixSize(4)                        # shrinking runs ensureTypes
    assert ph.particles.type.max() < 4
    ph.update(2)
    ph.setParticleCount(1000)
    ph.setPositions()
    assert (ph.particles.velocity == 0).all()
    ph.update(3)
    snap = ph.snapshot()
    assert snap.positions.shape == (3000,) and snap.particleCount == 1000
    assert snap.typeCount.sum() == 1000 and len(snap.typeCount) == 4
    ph.kill()


def test_save_load_round_trip_steps_identically():
    from particle_life import saves

    print("\n[physics] case=save_load", flush=True)
    a = _physics(4000, seed=9)
    a.settings.dt = 0.02
    a.update(3)
    buf = io.BytesIO()
    saves.save_physics(a, buf)
    buf.seek(0)
    b = _physics(10, seed=10)
    b.settings.dt = 0.02
    saves.load_physics(b, buf)
    np.testing.assert_array_equal(b.particles.position, a.particles.position)
    np.testing.assert_array_equal(b.particles.velocity, a.particles.velocity)
    assert b.settings.matrix == a.settings.matrix and b.settings.rmax == a.settings.rmax
    a.update(2)
    b.update(2)
    pa, pb = a.particles, b.particles
    np.testing.assert_array_equal(pa.position, pb.position)
    np.testing.assert_array_equal(pa.velocity, pb.velocity)
    a.kill()
    b.kill()


def test_matrix_shrunk_below_types_blocks_steps_until_fixed():
    """plife_set_settings may shrink the matrix below the particles' types (as
    Physics.setMatrixSize does before ensureTypes); stepping is refused with a
    state error until particles within the new size are set."""
    import numpy as np
    from particle_life import Engine, worlds
    from particle_life._native import PlifeError

    print("\n[physics] matrix shrunk below types", flush=True)
    w = worlds.uniform_world(401, 2_000, 6, rmax=0.03, dt=0.01)
    e = Engine(0)
    try:
        e.load_world(w)
        e.step(1)
        small = w.matrix[:3, :3].copy()
        e.configure(matrix=small, rmax=w.rmax, friction=w.friction, dt=w.dt, beta=w.beta,
                    force=w.force, wrap=w.wrap, dim=w.dim)
        with pytest.raises(PlifeError):
            e.step(1)
        with pytest.raises(PlifeError):
            e.set_types("rotate", 1)
        st = e.state()
        e.set_state(st["pos"], st["vel"], np.minimum(st["types"], 2))
        e.step(2)
        e.set_types("flip", 2)
        assert e.state()["types"].max() <= 2
    finally:
        e.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_slot_edits_step_like_a_full_upload(dim):
    """plife_set_particle_slots changes only the edited rows (ids kept), and the
    edited state steps bit-identically to the same state uploaded whole."""
    import numpy as np
    from particle_life import Engine, worlds
    from particle_life._native import PlifeError

    print(f"\n[physics] slot edits dim={dim}", flush=True)
    w = worlds.uniform_world(410 + dim, 6_000, 5, rmax=0.03 if dim == 2 else 0.06, dt=0.01, dim=dim)
    rng = np.random.default_rng(7)
    a, b = Engine(0), Engine(0)
    try:
        a.load_world(w)
        a.step(3)
        s0 = a.state()
        slots = rng.choice(w.n, 64, replace=False).astype(np.int32)
        pos = rng.random((64, 3)) * 1.2 - 0.1          # some outside the box: wrapped
        if dim == 2:
            pos[:, 2] = 0.0
        vel = (rng.random((64, 3)) - 0.5) * 0.01
        if dim == 2:
            vel[:, 2] = 0.0
        typ = rng.integers(0, 5, 64).astype(np.int32)
        a.set_slots(slots, pos, vel, typ)
        s1 = a.state()
        keep = np.ones(w.n, bool)
        keep[slots] = False
        for k in ("pos", "vel", "types", "ids"):
            np.testing.assert_array_equal(s1[k][keep], s0[k][keep])
        np.testing.assert_array_equal(s1["ids"], s0["ids"])
        np.testing.assert_array_equal(s1["types"][slots], typ)
        np.testing.assert_array_equal(s1["vel"][slots], vel.astype(np.float32).astype(np.float64))
        want = np.round((pos - np.floor(pos)) * 2.0**32) % 2.0**32 * 2.0**-32
        np.testing.assert_array_equal(s1["pos"][slots], want)
        b.configure_world(w)
        b.set_state(s1["pos"], s1["vel"], s1["types"])
        a.step(3)
        b.step(3)
        ra, rb = a.state(), b.state()
        for k in ("pos", "vel", "types"):
            np.testing.assert_array_equal(ra[k], rb[k])
        np.testing.assert_array_equal(s1["ids"][rb["ids"]], ra["ids"])
        for bad in ([0, 0], [w.n], [-1]):
            with pytest.raises(PlifeError):
                a.set_slots(bad, np.zeros((len(bad), 3)), np.zeros((len(bad), 3)),
                            np.zeros(len(bad), np.int32))
        with pytest.raises(PlifeError):
            a.set_slots([1], np.zeros((1, 3)), np.zeros((1, 3)), [5])
    finally:
        a.close()
        b.close()


def test_batch_slot_edits_and_gpu_physics_edits():
    import numpy as np
    from particle_life import Batch, GpuPhysics, ParticleLifeAccelerator, worlds

    print("\n[physics] batch slot edits, editParticles", flush=True)
    ws = [worlds.uniform_world(420 + k, 3_000, 4, rmax=0.03, dt=0.01) for k in range(3)]
    bt = Batch(0)
    try:
        ids = [bt.add_world(w) for w in ws]
        bt.step(2)
        bt.sync()
        before = [bt.state(i) for i in ids]
        bt.set_slots(ids[1], [5, 17], [[0.25, 0.5, 0], [0.75, 0.125, 0]], np.zeros((2, 3)), [3, 0])
        after = [bt.state(i) for i in ids]
        for k in (0, 2):
            for key in ("pos", "vel", "types", "ids"):
                np.testing.assert_array_equal(after[k][key], before[k][key])
        np.testing.assert_array_equal(after[1]["pos"][[5, 17]], [[0.25, 0.5, 0], [0.75, 0.125, 0]])
        np.testing.assert_array_equal(after[1]["types"][[5, 17]], [3, 0])
    finally:
        bt.close()
    phys = GpuPhysics(ParticleLifeAccelerator(0.3), particleCount=2_000)
    try:
        phys.update(2)
        phys.snapshot()
        phys.editParticles([0, 1], position=[[0.5, 0.5, 0], [1.25, -0.5, 0]], type_=[1, 2])
        snap = phys.snapshot()
        np.testing.assert_array_equal(snap.positions[:6], [0.5, 0.5, 0, 0.25, 0.5, 0])
        st = phys._engine.state()
        np.testing.assert_array_equal(st["pos"][:2], [[0.5, 0.5, 0], [0.25, 0.5, 0]])
        np.testing.assert_array_equal(st["types"][:2], [1, 2])
        phys.update(2)
    finally:
        phys.kill()


@pytest.mark.parametrize("dim,n", [(2, 300_001), (3, 140_000), (2, 70_000)])
def test_chunked_download_returns_the_uploaded_state(dim, n):
    """Downloads of 2 x 65,536 particles or more move in up to 8 chunks that
    host threads convert while later chunks are still in flight.  Every array,
    requested alone or with the others, into fresh or reused arrays, equals
    the uploaded state: positions on the 2^-24 grid and FP32-representable
    velocities pass through the device exactly.  Also for a world that is not
    the first of a batch (its slots start past the other world's)."""
    from particle_life import Batch, Engine, worlds

    print(f"\n[physics] case=chunked_download dim={dim} n={n}", flush=True)
    w = worlds.uniform_world(31 + dim, n, 5, rmax=0.01, dt=0.01, dim=dim)
    rng = np.random.default_rng(dim)
    w.vel = rng.uniform(-1e-3, 1e-3, (n, 3)).astype(np.float32).astype(np.float64)
    if dim == 2:
        w.vel[:, 2] = 0.0

    def check_against_upload(st):
        np.testing.assert_array_equal(np.sort(st["ids"]), np.arange(n))
        o = np.argsort(st["ids"])
        np.testing.assert_array_equal(st["pos"][o], w.pos)
        np.testing.assert_array_equal(st["vel"][o], w.vel)
        np.testing.assert_array_equal(st["types"][o], w.types)

    e = Engine(0)
    e.load_world(w)
    full = e.state()
    check_against_upload(full)
    pos, vel = np.full((n, 3), np.nan), np.full((n, 3), np.nan)
    typ, ids = np.full(n, -1, np.int32), np.full(n, -1, np.int32)
    e.download(pos=pos)
    np.testing.assert_array_equal(pos, full["pos"])
    e.download(vel=vel)
    np.testing.assert_array_equal(vel, full["vel"])
    e.download(types=typ)
    np.testing.assert_array_equal(typ, full["types"])
    e.download(ids=ids)
    np.testing.assert_array_equal(ids, full["ids"])
    # the same arrays again after steps, against a fresh full download
    e.step(2)
    e.download(pos, vel, typ)
    after = e.state()
    np.testing.assert_array_equal(pos, after["pos"])
    np.testing.assert_array_equal(vel, after["vel"])
    np.testing.assert_array_equal(typ, after["types"])
    with pytest.raises(ValueError):
        e.download(pos=np.empty((n, 3), np.float32))
    e.close()

    with Batch(0) as b:
        small = worlds.uniform_world(7, 1000, 3, rmax=0.05, dt=0.01, dim=dim)
        b.add_world(small)
        wid = b.add_world(w)
        check_against_upload(b.state(wid))


def test_snapshot_take_reuses_its_arrays():
    """PhysicsSnapshot.take rewrites the snapshot's arrays in place while the
    particle count stays (app/PhysicsSnapshot.java:39-44), new ones when it
    changes; the values are those of the physics' particles."""
    from particle_life import PhysicsSnapshot

    print("\n[physics] case=snapshot_take", flush=True)
    ph = _physics(2000)
    ph.update(2)
    s = PhysicsSnapshot()
    s.take(ph)
    arrays = (s.positions, s.velocities, s.types)
    ph.update(1)
    s.take(ph)
    assert all(a is b for a, b in zip(arrays, (s.positions, s.velocities, s.types)))
    p = ph.particles                              # host view of the same state
    np.testing.assert_array_equal(s.positions, p.position.reshape(-1))
    np.testing.assert_array_equal(s.velocities, p.velocity.reshape(-1))
    np.testing.assert_array_equal(s.types, p.type)
    assert s.particleCount == 2000 and s.typeCount.sum() == 2000 and s.snapshotTime > 0
    ph.setParticleCount(1500)
    s.take(ph)                                    # from the host view (not yet uploaded)
    assert s.positions.shape == (4500,) and s.positions is not arrays[0]
    np.testing.assert_array_equal(s.types, ph.particles.type)
    ph.kill()
