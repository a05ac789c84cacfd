"""Batches of worlds on the device: every world of a batch must match the
same world stepped alone (bit-identical: same kernels, same candidate order)
and the FP64 oracle (tolerances of parity_util), through joins and leaves.
Every case prints its name first so a failing GPU run names it."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from parity_util import VEL_ABS, VEL_REL, oracle_settings

pytestmark = pytest.mark.gpu


def _mixed_worlds(dim=2):
    from particle_life import worlds

    w = worlds.uniform_world
    out = [
        w(201, 3_000, 3, rmax=0.05, dt=0.01, beta=0.25, dim=dim),
        w(202, 20_000, 12, rmax=0.02, dt=0.01, beta=0.35, dim=dim),
        w(203, 1, 4, rmax=0.03, dt=0.01, dim=dim),                         # lone particle
        w(204, 0, 5, rmax=0.03, dt=0.01, dim=dim),                         # empty world
        w(205, 6_000, 6, rmax=0.04, dt=0.02, beta=0.3, wrap=False, dim=dim),   # clamped
        w(206, 800, 7, rmax=0.4, dt=0.01, dim=dim),                        # nx = 2 (literal stencil)
        worlds.gaussian_blob_world(207, 9_000, 5, rmax=0.03, dt=0.01, beta=0.3, nblobs=3,
                                   sigma=0.02, dim=dim),
        w(208, 4_000, 40, rmax=0.03, dt=0.01, dim=dim),   # > 32 types: the batch's matrices in global memory
    ]
    return out


def _alone(w, nsteps, **tune):
    from particle_life import Engine

    e = Engine(0)
    if tune:
        e.set_tuning(**tune)
    e.load_world(w)
    if nsteps:
        e.step(nsteps)
    st = e.state()
    pairs = e.pair_count()
    e.close()
    return st, pairs


def _assert_same(a, b, what):
    for k in ("pos", "vel", "ids", "types"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what}: {k}")


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_worlds_alone(dim):
    """5 steps (two graph replays + one direct step) of a mixed batch equal
    each world stepped alone, bit for bit, and per-world pair counts agree."""
    from particle_life import Batch

    print(f"\n[batch] case=mixed_alone dim={dim}", flush=True)
    ws = _mixed_worlds(dim)
    b = Batch(0)
    ids = [b.add_world(w) for w in ws]
    b.step(5)
    for wid, w in zip(ids, ws):
        st, pairs = _alone(w, 5)
        _assert_same(b.state(wid), st, f"world {wid}")
        info = b.info(wid)
        assert info["steps"] == 5 and info["n"] == w.n
        assert info["pairs"] == pairs, (wid, info["pairs"], pairs)
    b.close()


def test_batch_one_step_oracle():
    """Each world's step in the batch against the FP64 oracle on its own state."""
    from particle_life import Batch

    print("\n[batch] case=oracle_per_world", flush=True)
    ws = _mixed_worlds(2)
    b = Batch(0)
    ids = [b.add_world(w) for w in ws]
    b.step(2)
    before = {wid: b.state(wid) for wid in ids}
    b.step(1)
    for wid, w in zip(ids, ws):
        cur = before[wid]
        ref = oracle.step(oracle_settings(w), cur["pos"], cur["vel"], cur["types"], ids=cur["ids"],
                          nthreads=8, debug=True)
        new = b.state(wid)
        np.testing.assert_array_equal(new["ids"], ref["ids"], err_msg=f"world {wid}: order")
        assert b.info(wid)["pairs"] == ref["pair_count"]
        dv = np.abs(new["vel"] - ref["vel"])
        assert not (dv > VEL_ABS + VEL_REL * np.abs(ref["vel"])).any(), f"world {wid}: velocity"
    b.close()


def test_batch_join_and_leave():
    """Worlds leave and join between updates; the survivors' trajectories are
    unchanged (equal to running alone) and newcomers start from their own state."""
    from particle_life import Batch, worlds

    print("\n[batch] case=join_leave", flush=True)
    ws = _mixed_worlds(2)
    extra = worlds.uniform_world(208, 12_000, 8, rmax=0.025, dt=0.01, beta=0.3)
    b = Batch(0)
    ids = [b.add_world(w) for w in ws]
    b.step(3)
    b.remove(ids[1])                     # a large world in the middle
    b.remove(ids[3])                     # the empty one
    new_id = b.add_world(extra)
    b.step(4)
    assert b.count() == (len(ws) - 1, sum(w.n for w in ws) - ws[1].n + extra.n)
    for k, (wid, w) in enumerate(zip(ids, ws)):
        if k in (1, 3):
            continue
        st, _ = _alone(w, 7)
        _assert_same(b.state(wid), st, f"world {wid}")
        assert b.info(wid)["steps"] == 7
    st, pairs = _alone(extra, 4)
    _assert_same(b.state(new_id), st, "newcomer")
    assert b.info(new_id)["pairs"] == pairs
    with pytest.raises(ValueError):
        b.state(ids[1])
    b.close()


def test_batch_live_settings_and_tuning():
    """Per-world settings change between updates; launch tuning does not change
    results."""
    from particle_life import Batch

    print("\n[batch] case=settings_tuning", flush=True)
    ws = _mixed_worlds(2)[:3]
    res = []
    for tune in (dict(), dict(block_homes=64, list_cap=8, stage_slack=0.05, direct_launch=True)):
        b = Batch(0)
        b.set_tuning(**tune)
        ids = [b.add_world(w) for w in ws]
        b.step(2)
        w1 = ws[1]
        b.set_world_settings(ids[1], matrix=-w1.matrix, rmax=0.03, friction=w1.friction, dt=0.005,
                             beta=0.4, force=2.0, wrap=True, dim=2)
        b.step(3)
        res.append([b.state(i) for i in ids])
        b.close()
    for a, c in zip(*res):
        _assert_same(a, c, "tuning")


def test_work_balanced_launches_equal_worlds_alone():
    """A clustered world among uniform ones makes the block table's work
    skewed, so later updates dispatch its blocks heaviest first (and across
    layout changes, with the work carried over); every world still equals the
    same world stepped alone in one update."""
    from particle_life import Batch, Engine, worlds

    print("\n[batch] case=work_balanced", flush=True)
    ws = [worlds.gaussian_blob_world(520, 40_000, 6, rmax=0.02, dt=0.01, beta=0.3, nblobs=4,
                                     sigma=0.01, dim=2)]
    ws += [worlds.uniform_world(521 + k, 20_000, 5, rmax=0.015, dt=0.01) for k in range(4)]
    b = Batch(0)
    try:
        ids = [b.add_world(w) for w in ws]
        for _ in range(3):
            b.step(1)
            b.sync()
        extra = b.add_world(worlds.uniform_world(530, 5_000, 4, rmax=0.02, dt=0.01))
        for _ in range(3):
            b.step(1)
            b.sync()
        b.remove(extra)
        b.step(1)
        b.sync()
        got = [b.state(i) for i in ids]
    finally:
        b.close()
    e = Engine(0)
    try:
        for w, g in zip(ws, got):
            e.load_world(w)
            e.step(7)
            s = e.state()
            for key in ("pos", "vel", "types", "ids"):
                np.testing.assert_array_equal(g[key], s[key], err_msg=key)
    finally:
        e.close()


def test_many_tiny_worlds():
    """1,200 tiny worlds (0-60 particles, 1-4 types, rmax up to 0.5 so some
    grids have fewer than 3 cells per axis) in one batch: a sample of them
    equals the same world stepped alone, and pair counts add up."""
    from particle_life import Batch, Engine, worlds

    print("\n[batch] case=many_tiny_worlds", flush=True)
    rng = np.random.default_rng(11)
    ws = []
    for k in range(1200):
        n = int(rng.choice([0, 1, 2, 7, 60]))
        ws.append(worlds.uniform_world(600 + k, n, int(rng.integers(1, 5)),
                                       rmax=float(rng.uniform(0.05, 0.5)), dt=0.01,
                                       wrap=bool(k % 3)))
    b = Batch(0)
    try:
        ids = [b.add_world(w) for w in ws]
        for _ in range(3):
            b.step(1)
            b.sync()
        sample = list(range(0, 1200, 41))
        got = {k: b.state(ids[k]) for k in sample}
        info = {k: b.info(ids[k]) for k in sample}
    finally:
        b.close()
    e = Engine(0)
    try:
        for k in sample:
            e.load_world(ws[k])
            e.step(3)
            s = e.state()
            for key in ("pos", "vel", "types", "ids"):
                np.testing.assert_array_equal(got[k][key], s[key], err_msg=f"world {k}: {key}")
            assert info[k]["pairs"] == e.pair_count()
    finally:
        e.close()


def test_full_size_batch_one_step_oracle():
    """The 512-world batch at full size (~7.9M particles, 3-12 types, rmax and
    beta per world): after two steps, one more step of every world against
    the oracle on its own state -- order, exact pair count, velocities and
    positions within tolerance."""
    from particle_life import Batch, worlds

    print("\n[batch] case=c4_full_size", flush=True)
    ws = worlds.config4()
    b = Batch(0)
    ids = [b.add_world(w) for w in ws]
    b.step(2)
    before = {wid: b.state(wid) for wid in ids}
    b.step(1)
    total = 0
    for wid, w in zip(ids, ws):
        cur = before[wid]
        ref = oracle.step(oracle_settings(w), cur["pos"], cur["vel"], cur["types"], ids=cur["ids"],
                          nthreads=16, debug=True)
        new = b.state(wid)
        np.testing.assert_array_equal(new["ids"], ref["ids"], err_msg=f"world {wid}: order")
        assert b.info(wid)["pairs"] == ref["pair_count"], f"world {wid}: pairs"
        dv = np.abs(new["vel"] - ref["vel"])
        assert not (dv > VEL_ABS + VEL_REL * np.abs(ref["vel"])).any(), f"world {wid}: velocity"
        dp = np.abs(new["pos"] - ref["pos"])
        dp = np.minimum(dp, 1.0 - dp)
        assert not (dp > 2.0 ** -32 + w.dt * (VEL_ABS + VEL_REL * np.abs(ref["vel"]))).any(), \
            f"world {wid}: position"
        total += ref["pair_count"]
    print(f"  worlds {len(ws)}, particles {sum(w.n for w in ws)}, pairs {total}", flush=True)
    b.close()
