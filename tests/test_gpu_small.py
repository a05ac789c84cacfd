"""The one-launch step of worlds that fit one workgroup's LDS (k_step_small)
against the multi-launch path (bin, scans, radix sort, force kernels) on the
same inputs: states, pair counts, the sort's by-products (cell keys, cell
ranges, permutation) and neighbour sets must be bit-identical, alone and in
batches, through graph replays and direct launches.  The oracle cases of
test_gpu_parity.py that fit (C1 among them) check the one-launch step against
the FP64 oracle, since it is the automatic choice for them."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = {
    # name: (seed, n, types, rmax, dt, extra)
    "c1_10k_2d": (101, 10_000, 6, 0.02, 0.02, {}),
    "clamped_2d": (102, 6_000, 5, 0.03, 0.02, {"wrap": False}),
    "noninteger_cells_2d": (103, 8_000, 7, 0.07, 0.02, {}),
    "nx3_2d": (104, 2_000, 4, 0.30, 0.02, {}),
    "nx2_repeated_cells_2d": (105, 400, 3, 0.40, 0.02, {}),
    "nx1_2d": (106, 50, 3, 0.9, 0.02, {}),
    "dense_2d": (112, 9_000, 6, 0.2, 0.01, {}),                # 360 per cell
    "sparse_cells_2d": (113, 3_000, 5, 0.008, 0.02, {}),       # 15,625 cells, 0.2 per cell
    "3d": (114, 6_000, 6, 0.1, 0.01, {"dim": 3}),
    "3d_clamped": (110, 8_000, 5, 0.09, 0.01, {"dim": 3, "wrap": False}),
    "3d_nx2": (111, 300, 3, 0.45, 0.01, {"dim": 3}),
    "many_types_2d": (140, 8_000, 40, 0.03, 0.02, {}),
    "one_particle_2d": (115, 1, 3, 0.05, 0.02, {}),
}


def _world(seed, n, ntypes, rmax, dt, **kw):
    from particle_life import worlds

    return worlds.uniform_world(seed, n, ntypes, rmax=rmax, dt=dt, **kw)


def _run(w, steps, **tune):
    from particle_life import Engine

    e = Engine(0)
    try:
        e.set_tuning(**tune)
        e.set_debug(True)
        e.load_world(w)
        e.step(steps)
        st = e.state()
        out = {k: st[k] for k in ("pos", "vel", "ids", "types")}
        out["pairs"] = e.pair_count()
        c = e.cells()
        out.update(keys=c["keys"], cell_end=c["cell_end"], perm=c["perm"])
        out["nbr_count"], out["nbr_hash"] = e.neighbor_summary()
        tn = e.tuning()
        out["mode"], out["homes"] = tn["small_step"], tn["block_homes"]
    finally:
        e.close()
    return out


def _assert_same(a, b, what):
    for k in ("pos", "vel", "ids", "types", "keys", "cell_end", "perm", "nbr_count", "nbr_hash"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")
    assert a["pairs"] == b["pairs"], (what, a["pairs"], b["pairs"])


@pytest.mark.parametrize("name", list(SMALL))
def test_one_launch_equals_multi_launch(name):
    """5 steps (two graph replays and a direct step) on each path; and the
    one-launch path with every step launched directly."""
    seed, n, t, rmax, dt, kw = SMALL[name]
    print(f"\n[small] case={name} n={n} rmax={rmax}", flush=True)
    w = _world(seed, n, t, rmax, dt, **kw)
    multi = _run(w, 5, small_step=1)
    one = _run(w, 5, small_step=2)
    assert multi["mode"] == 1 and one["mode"] == 2, (multi["mode"], one["mode"])
    _assert_same(one, multi, name)
    direct = _run(w, 5, small_step=2, direct_launch=True)
    _assert_same(direct, multi, name + " direct")


@pytest.mark.parametrize("homes", [64, 128, 256])
def test_homes_per_workgroup(homes):
    """Any number of homes per workgroup gives the same bits (each workgroup
    sorts the whole world and steps its own run of homes)."""
    print(f"\n[small] case=homes_{homes}", flush=True)
    w = _world(116, 7_000, 6, 0.025, 0.02)
    ref = _run(w, 3, small_step=1)
    got = _run(w, 3, small_step=2, block_homes=homes)
    assert got["mode"] == 2
    _assert_same(got, ref, f"homes {homes}")


def test_automatic_choice():
    """Automatic tuning times both paths on the first update of at least 24
    steps of a layout that fits the one-launch step and keeps the faster
    (either way the bits of the multi-launch path); a world that does not fit,
    or another launch-shape field set, takes the multi-launch path."""
    print("\n[small] case=automatic_choice", flush=True)
    w = _world(117, 5_000, 6, 0.03, 0.02)
    auto = _run(w, 26)
    assert auto["mode"] in (1, 2)
    _assert_same(auto, _run(w, 26, small_step=1), "automatic")
    assert _run(_world(118, 30_000, 6, 0.02, 0.02), 8)["mode"] == 1
    assert _run(w, 8, list_cap=16)["mode"] == 1


@pytest.mark.parametrize("dim", [2, 3])
def test_automatic_homes_per_workgroup(dim):
    """Where block_homes is automatic, the first update of at least 24 steps
    of a layout on the multi-launch path times 256 and 128 homes per force
    workgroup and keeps the faster; either way the bits are those of a fixed
    launch shape, and a second world of the same class reuses the choice."""
    from particle_life import Engine

    print(f"\n[small] case=automatic_homes dim={dim}", flush=True)
    n, rmax = (40_000, 0.02) if dim == 2 else (30_000, 0.06)     # too large for the one-launch step
    w = _world(119, n, 6, rmax, 0.01, dim=dim)
    auto = _run(w, 26)
    assert auto["mode"] == 1 and auto["homes"] in (128, 256), (auto["mode"], auto["homes"])
    _assert_same(auto, _run(w, 26, block_homes=256), "automatic homes")
    _assert_same(auto, _run(w, 26, block_homes=128), "automatic homes vs 128")
    e = Engine(0)
    try:
        e.load_world(w)
        e.step(26)
        first = e.tuning()["block_homes"]
        e.load_world(_world(120, n, 6, rmax, 0.01, dim=dim))
        e.step(3)                                    # too short to measure: the class's choice
        assert e.tuning()["block_homes"] == first
    finally:
        e.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_of_small_worlds(dim):
    """A batch whose worlds all fit steps in one launch (one ticket counter
    over all workgroups); each world equals the same world stepped alone on the
    multi-launch path, and per-world pair counts agree.  A world that leaves
    and one that joins change the launch between updates."""
    from particle_life import Batch, Engine, worlds

    print(f"\n[small] case=batch dim={dim}", flush=True)
    u = worlds.uniform_world
    r = (lambda r2, r3: r2 if dim == 2 else r3)        # 3D grids of at most 20^3 cells
    ws = [u(301, 3_000, 3, rmax=0.05, dt=0.01, beta=0.25, dim=dim),
          u(302, 1, 4, rmax=r(0.03, 0.07), dt=0.01, dim=dim),
          u(303, 0, 5, rmax=r(0.03, 0.07), dt=0.01, dim=dim),
          u(304, 5_000, 6, rmax=r(0.04, 0.1), dt=0.02, wrap=False, dim=dim),
          u(305, 800, 7, rmax=0.4, dt=0.01, dim=dim),
          u(306, 4_000, 40, rmax=r(0.03, 0.1), dt=0.01, dim=dim)]
    b = Batch(0)
    try:
        b.set_tuning(small_step=2)
        ids = [b.add_world(w) for w in ws]
        b.step(4)
        assert b.tuning()["small_step"] == 2
        for w, wid in zip(ws, ids):
            e = Engine(0)
            e.set_tuning(small_step=1)
            e.load_world(w)
            e.step(4)
            st, pairs = e.state(), e.pair_count()
            e.close()
            got = b.state(wid)
            for k in ("pos", "vel", "ids", "types"):
                np.testing.assert_array_equal(got[k], st[k], err_msg=f"{w.name}: {k}")
            assert b.info(wid)["pairs"] == pairs, (w.name, b.info(wid)["pairs"], pairs)
        # leave and join: the next update re-plans the launch
        b.remove(ids[0])
        late = u(307, 2_500, 5, rmax=0.05, dt=0.01, dim=dim)
        wid = b.add_world(late)
        b.step(3)
        e = Engine(0)
        e.set_tuning(small_step=1)
        e.load_world(late)
        e.step(3)
        st = e.state()
        e.close()
        got = b.state(wid)
        for k in ("pos", "vel", "ids", "types"):
            np.testing.assert_array_equal(got[k], st[k], err_msg=f"late world: {k}")
    finally:
        b.close()


def _sequence(w, small, updates, edit=None):
    from particle_life import Engine

    e = Engine(0)
    try:
        e.set_tuning(small_step=small)
        e.load_world(w)
        for k, u in enumerate(updates):
            e.step(u)
            if edit is not None and k == 0:
                edit(e)
        st = e.state()
        return {"pos": st["pos"], "vel": st["vel"], "ids": st["ids"], "types": st["types"],
                "pairs": e.pair_count(), "mode": e.tuning()["small_step"]}
    finally:
        e.close()


def test_update_sequences_and_edits():
    """Updates of odd and even step counts (direct launches and graph replays
    starting from either sorted-key buffer) and particle edits between updates
    (the previous step's keys no longer describe the state) give the
    multi-launch path's bits."""
    print("\n[small] case=update_sequences", flush=True)
    w = _world(119, 6_000, 5, 0.03, 0.02)
    ups = [1, 5, 2, 3, 4, 7]
    a = _sequence(w, 1, ups)
    b = _sequence(w, 2, ups)
    assert b["mode"] == 2
    for k in ("pos", "vel", "ids", "types"):
        np.testing.assert_array_equal(b[k], a[k], err_msg=k)
    assert a["pairs"] == b["pairs"]

    rng = np.random.default_rng(7)
    slots = np.sort(rng.choice(w.n, 1_500, replace=False)).astype(np.int32)
    pos = rng.random((slots.size, 3)) * 2.0 - 1.0
    pos[:, 2] = 0.0
    vel = np.zeros((slots.size, 3))
    typ = rng.integers(0, 5, slots.size).astype(np.int32)

    def edit(e):
        e.set_slots(slots, pos, vel, typ)

    a = _sequence(w, 1, [4, 3, 2], edit)
    b = _sequence(w, 2, [4, 3, 2], edit)
    for k in ("pos", "vel", "ids", "types"):
        np.testing.assert_array_equal(b[k], a[k], err_msg="edited " + k)
    assert a["pairs"] == b["pairs"]
