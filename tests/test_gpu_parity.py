"""Device step vs the FP64 oracle on identical states (see parity_util for the
tolerances).  Every case prints its name first so a failing GPU run names it."""
from __future__ import annotations

import numpy as np
import pytest

from parity_util import check_one_step, free_run_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from particle_life import Engine

    e = Engine(0)
    yield e
    e.close()


def _world(seed, n, ntypes, rmax, dt, **kw):
    from particle_life import worlds

    return worlds.uniform_world(seed, n, ntypes, rmax=rmax, dt=dt, **kw)


CASES = {
    # name: (seed, n, types, rmax, dt, extra)
    "c1_10k_2d": (101, 10_000, 6, 0.02, 0.02, {}),
    "clamped_2d": (102, 6_000, 5, 0.03, 0.02, {"wrap": False}),
    "noninteger_cells_2d": (103, 8_000, 7, 0.07, 0.02, {}),          # last cell wider (Q2)
    "nx3_2d": (104, 2_000, 4, 0.30, 0.02, {}),
    "nx2_repeated_cells_2d": (105, 400, 3, 0.40, 0.02, {}),          # Q1: repeated stencil cells
    "nx1_2d": (106, 50, 3, 0.9, 0.02, {}),
    "strong_force_2d": (107, 5_000, 8, 0.025, 0.05, {"force": 40.0, "beta": 0.2}),
    "dense_2d": (108, 40_000, 6, 0.05, 0.01, {}),                    # 100/cell: wide strips
    "3d": (109, 20_000, 6, 0.06, 0.01, {"dim": 3}),
    "3d_clamped": (110, 8_000, 5, 0.09, 0.01, {"dim": 3, "wrap": False}),
    "3d_nx2": (111, 300, 3, 0.45, 0.01, {"dim": 3}),
    # matrices too large for LDS (> 32 types): read from global memory
    "many_types_2d": (140, 8_000, 40, 0.03, 0.02, {}),
    "max_types_2d": (141, 20_000, 256, 0.02, 0.02, {}),
    "many_types_3d": (142, 20_000, 64, 0.06, 0.01, {"dim": 3}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_one_step_parity(eng, name):
    seed, n, t, rmax, dt, kw = CASES[name]
    print(f"\n[parity] case={name} n={n} rmax={rmax}", flush=True)
    w = _world(seed, n, t, rmax, dt, **kw)
    eng.load_world(w)
    for k in range(3):
        st = check_one_step(eng, w)
        print(f"  step {k}: {st}", flush=True)


def test_ten_steps_each_checked(eng):
    """Ten consecutive steps, each compared with the oracle stepping the device state."""
    from particle_life import worlds

    w = worlds.config1()
    print("\n[parity] case=c1_ten_steps", flush=True)
    eng.load_world(w)
    for k in range(10):
        check_one_step(eng, w, debug=(k % 3 == 0))


def test_trajectory_10_steps(eng):
    """Free-running 10-step trajectory stays within 1e-8 (unit box) of the oracle's."""
    from particle_life import worlds

    w = worlds.config1()
    print("\n[parity] case=c1_trajectory", flush=True)
    err = free_run_error(eng, w, 10)
    print(f"  max position error after 10 steps: {err:.3e}")
    assert err < 1e-8


def test_heavy_cells_global_path(eng):
    """Two dense blobs (staged in several windows) in a sparse world whose
    workgroups span more grid rows than the staged layout holds (global path)."""
    from particle_life import worlds

    print("\n[parity] case=heavy_blob_2d", flush=True)
    w = worlds.gaussian_blob_world(112, 20_000, 6, rmax=0.02, dt=0.01, beta=0.3, nblobs=2,
                                   sigma=0.02, dim=2)
    eng.load_world(w)
    for _ in range(2):
        check_one_step(eng, w)


def test_clustered_3d(eng):
    from particle_life import worlds

    print("\n[parity] case=clustered_3d", flush=True)
    w = worlds.gaussian_blob_world(113, 40_000, 12, rmax=0.04, dt=0.01, beta=0.25, nblobs=8,
                                   sigma=0.03, dim=3)
    eng.load_world(w)
    check_one_step(eng, w)


DENSE = {
    # name: (seed, n, types, rmax, dt, beta, blobs, sigma, dim) -- max cell occupancy
    "blob_3d_4k_per_cell": (120, 30_000, 12, 0.02, 0.01, 0.25, 4, 0.006, 3),    # 4284
    "blob_2d_7k_per_cell": (121, 30_000, 6, 0.02, 0.01, 0.3, 3, 0.006, 2),      # 7364
    "blob_3d_c3_density": (122, 120_000, 12, 0.01, 0.01, 0.25, 8, 0.008, 3),    # 1392
}


@pytest.mark.parametrize("name", list(DENSE))
def test_dense_neighbourhoods_windowed(eng, name):
    """Neighbourhoods of thousands of particles per cell: the staged
    neighbourhood exceeds the LDS buffer and is scanned in several windows."""
    from particle_life import worlds

    seed, n, t, rmax, dt, beta, nb, sig, dim = DENSE[name]
    print(f"\n[parity] case={name} n={n}", flush=True)
    w = worlds.gaussian_blob_world(seed, n, t, rmax=rmax, dt=dt, beta=beta, nblobs=nb,
                                   sigma=sig, dim=dim)
    eng.load_world(w)
    st = check_one_step(eng, w)
    print(f"  {st}", flush=True)


@pytest.mark.parametrize("dim", [2, 3])
def test_tiny_staging_buffer(eng, dim):
    """A staging buffer of a few dozen particles forces every workgroup through
    many windows, and 8-entry candidate lists through many partial drains;
    results must stay within tolerance of the oracle."""
    print(f"\n[parity] case=tiny_staging_{dim}d", flush=True)
    w = _world(123 + dim, 20_000, 6, 0.03 if dim == 2 else 0.07, 0.01, dim=dim)
    eng.set_tuning(stage_slack=0.01, list_cap=8, block_homes=64)
    try:
        eng.load_world(w)
        for _ in range(2):
            check_one_step(eng, w)
    finally:
        eng.set_tuning()


def test_launch_plans_bit_identical(eng):
    """Every home visits its candidates in the same order whatever the
    workgroup size, list capacity and staging windows, so states after several
    steps are bit-identical across launch plans."""
    from particle_life import worlds

    print("\n[parity] case=plans_bit_identical", flush=True)
    w = worlds.gaussian_blob_world(126, 30_000, 6, rmax=0.03, dt=0.01, beta=0.3, nblobs=6,
                                   sigma=0.03, dim=2)
    plans = [dict(), dict(block_homes=64, list_cap=8, stage_slack=0.05),
             dict(block_homes=128, list_cap=64, stage_slack=4.0),
             dict(block_homes=256, list_cap=16, stage_slack=0.3, direct_launch=True),
             dict(sparse_worlds=2), dict(sparse_worlds=1, block_homes=64),
             dict(windowed_only=True), dict(windowed_only=True, block_homes=128)]
    ref = None
    for p in plans:
        eng.set_tuning(**p)
        eng.load_world(w)
        eng.step(4)
        st = eng.state()
        got = (st["pos"], st["vel"], st["ids"], eng.pair_count())
        if ref is None:
            ref = got
            continue
        for k in range(3):
            np.testing.assert_array_equal(got[k], ref[k], err_msg=f"plan {p}")
        assert got[3] == ref[3]
    eng.set_tuning()


@pytest.mark.parametrize("dim,wrap", [(2, True), (2, False), (3, True), (3, False)])
def test_balanced_path_equals_windowed(eng, dim, wrap):
    """Uniform worlds step on the balanced path (one staging window, homes
    dealt to lanes by their kept-candidate counts) bit-identically to the
    windowed path with per-thread lists, neighbour sets included; both
    against the oracle."""
    print(f"\n[parity] case=balanced_vs_windowed dim={dim} wrap={wrap}", flush=True)
    w = _world(160 + dim + 2 * wrap, 40_000 if dim == 2 else 30_000, 7,
               0.012 if dim == 2 else 0.05, 0.01, dim=dim, wrap=wrap)
    out = []
    try:
        for win in (False, True):
            eng.set_tuning(windowed_only=win)
            eng.load_world(w)
            check_one_step(eng, w)
            eng.step(3)
            check_one_step(eng, w)
            st = eng.state()
            cnt, hsh = eng.neighbor_summary()
            out.append((st["pos"], st["vel"], st["ids"], cnt, hsh))
    finally:
        eng.set_tuning()
    for k in range(5):
        np.testing.assert_array_equal(out[0][k], out[1][k])


SPARSE = {
    # name: (seed, n, types, rmax, dim, wrap) -- particles per cell
    "sparse_3d": (130, 30_000, 6, 0.02, 3, True),          # 0.24
    "sparse_3d_clamped": (131, 20_000, 5, 0.03, 3, False),  # 0.55
    "sparse_2d": (132, 3_000, 4, 0.01, 2, True),           # 0.30
    "sparse_2d_clamped": (133, 2_000, 4, 0.015, 2, False),  # 0.46
    "sparse_3d_48_types": (134, 20_000, 48, 0.02, 3, True),  # matrix in global memory
}


@pytest.mark.parametrize("name", list(SPARSE))
def test_sparse_world_kernel(eng, name):
    """Sparse worlds: the sparse-world kernel (one thread per particle,
    stencil rows read as runs of the cell-sorted order) matches the oracle,
    and states after several steps are bit-identical whichever kernel steps
    the world."""
    seed, n, t, rmax, dim, wrap = SPARSE[name]
    print(f"\n[parity] case={name} n={n}", flush=True)
    w = _world(seed, n, t, rmax, 0.01, dim=dim, wrap=wrap)
    eng.set_tuning(sparse_worlds=2)
    try:
        eng.load_world(w)
        st = check_one_step(eng, w)
        print(f"  {st}", flush=True)
        ref = None
        for mode in (2, 1, 0):
            eng.set_tuning(sparse_worlds=mode)
            eng.load_world(w)
            eng.step(5)
            s = eng.state()
            got = (s["pos"], s["vel"], s["ids"], eng.pair_count())
            if ref is None:
                ref = got
                continue
            for k in range(3):
                np.testing.assert_array_equal(got[k], ref[k], err_msg=f"sparse_worlds={mode}")
            assert got[3] == ref[3]
    finally:
        eng.set_tuning()


def test_planner_detects_clustering(eng):
    """The automatic planner switches a clustered world to the clustered launch
    shape once a step's clustering measure has reached the host, and a uniform
    world back; results stay within tolerance of the oracle throughout."""
    from particle_life import worlds

    print("\n[parity] case=planner_regimes", flush=True)
    eng.set_tuning()
    w = worlds.gaussian_blob_world(122, 120_000, 12, rmax=0.01, dt=0.01, beta=0.25, nblobs=8,
                                   sigma=0.008, dim=3)
    eng.load_world(w)
    eng.step(1)
    eng.sync()
    check_one_step(eng, w, debug=False)          # planned from the measured clustering
    assert eng.tuning()["block_homes"] == 128
    u = worlds.uniform_world(127, 50_000, 6, rmax=0.02, dt=0.01)
    eng.load_world(u)
    eng.step(1)
    eng.sync()
    check_one_step(eng, u, debug=False)
    assert eng.tuning()["block_homes"] == 256


def test_first_update_replans_after_one_step(eng):
    """A clustered 3D world's first long update runs its first step on the
    uniform plan, measures the clustering and steps the rest on the clustered
    plan, with the bits of a world stepped one update at a time."""
    from particle_life import Engine, worlds

    print("\n[parity] case=first_update_replans", flush=True)
    eng.set_tuning()
    w = worlds.gaussian_blob_world(128, 120_000, 12, rmax=0.01, dt=0.01, beta=0.25, nblobs=8,
                                   sigma=0.008, dim=3)
    eng.load_world(w)
    eng.step(4)
    assert eng.tuning()["block_homes"] == 128
    got = eng.state()
    with Engine(0) as e:
        e.load_world(w)
        for _ in range(4):
            e.step(1)
        ref = e.state()
    for k in ("pos", "vel", "ids", "types"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_edge_sizes(eng):
    """Empty world, one particle, two particles, coincident particles."""
    from particle_life import worlds

    print("\n[parity] case=edge_sizes", flush=True)
    base = _world(114, 3, 3, 0.05, 0.02)
    for n in (0, 1, 2):
        w = worlds.World(pos=base.pos[:n], vel=base.vel[:n], types=base.types[:n],
                         matrix=base.matrix, rmax=0.05, friction=base.friction, dt=0.02)
        eng.load_world(w)
        check_one_step(eng, w)
    # two coincident particles plus one within range: d^2 == 0 is excluded
    pos = np.array([[0.5, 0.5, 0], [0.5, 0.5, 0], [0.52, 0.5, 0]])
    w = worlds.World(pos=pos, vel=np.zeros((3, 3)), types=np.array([0, 1, 2], np.int32),
                     matrix=base.matrix, rmax=0.05, friction=base.friction, dt=0.02)
    eng.load_world(w)
    st = check_one_step(eng, w)
    assert st["pairs"] == 4


def test_seam_and_exact_distance(eng):
    """Pairs across the periodic seam and at exactly d = rmax (the predicate
    is <=, the force there is zero)."""
    from particle_life import worlds

    print("\n[parity] case=seam_and_ties", flush=True)
    r = 2.0 ** -6
    pos = np.array([
        [2.0 ** -20, 0.5, 0], [1 - 2.0 ** -12, 0.5, 0],       # across x seam
        [0.25, 2.0 ** -24, 0], [0.25, 1 - 2.0 ** -10, 0],      # across y seam
        [0.6, 0.6, 0], [0.6 + r, 0.6, 0],                      # exactly rmax apart
        [0.3, 0.3, 0], [0.3 + r, 0.3 + 2.0 ** -32, 0],         # just beyond
    ])
    n = len(pos)
    w = worlds.World(pos=pos, vel=np.zeros((n, 3)), types=np.arange(n, dtype=np.int32) % 3,
                     matrix=np.array([[0.5, -1, 1], [1, 0.25, -0.5], [-1, 1, 0.75]]),
                     rmax=r, friction=0.9, dt=0.02)
    eng.load_world(w)
    st = check_one_step(eng, w)
    assert st["pairs"] == 6


def test_dt_and_matrix_change_between_steps(eng):
    """Settings are read live each step (Physics.java:401-437; app/Main.java:292-295)."""
    w = _world(115, 5_000, 6, 0.03, 0.02)
    print("\n[parity] case=live_settings", flush=True)
    eng.load_world(w)
    check_one_step(eng, w)
    w.dt = 0.013
    w.matrix = -w.matrix
    w.rmax = 0.041
    eng.configure_world(w)
    check_one_step(eng, w)


def test_c2_first_ten_steps_each_checked(eng):
    """The 1M-particle 2D benchmark world (~78 interacting neighbours each):
    each of its first ten steps checked against the oracle stepping the same
    device state -- keys, ranges, permutation and neighbour sets exactly,
    velocities, accelerations and positions within tolerance."""
    from particle_life import worlds

    print("\n[parity] case=c2_full_size_ten_steps", flush=True)
    w = worlds.config2()
    eng.load_world(w)
    for k in range(10):
        st = check_one_step(eng, w, nthreads=16, debug=(k % 3 == 0))
        print(f"  step {k}: {st}", flush=True)
        assert 70 * w.n < st["pairs"] < 90 * w.n


def test_c3_checked_early_and_late(eng):
    """The 4M-particle clustered 3D world of 64 Gaussian blobs (~1,700
    neighbours each, windowed staging, the clustered launch plan once a step
    has measured the clustering), checked at step 3 and again at step 21,
    after the blobs have contracted."""
    from particle_life import worlds

    print("\n[parity] case=c3_full_size", flush=True)
    w = worlds.config3()
    eng.load_world(w)
    eng.step(2)                     # the planner sees the clustering after a step
    st = check_one_step(eng, w, nthreads=16, debug=True)
    print(f"  step 3: {st}", flush=True)
    assert 1000 * w.n < st["pairs"] < 2500 * w.n
    eng.step(17)
    st = check_one_step(eng, w, nthreads=16, debug=True)
    print(f"  step 21: {st}", flush=True)
    assert 1000 * w.n < st["pairs"] < 4000 * w.n


def test_single_particle_world_on_the_staged_kernel(eng):
    """One particle at slot 0 stepped by the staged force kernel (forced with
    sparse_worlds=1): a workgroup with a single home at the buffer's start
    reads no key before it."""
    print("\n[parity] case=single_particle_staged", flush=True)
    try:
        eng.set_tuning(sparse_worlds=1)
        for n in (1, 2, 3):
            w = _world(150 + n, n, 3, 0.05, 0.02)
            eng.load_world(w)
            for _ in range(2):
                st = check_one_step(eng, w)
                if n == 1:
                    assert st["pairs"] == 0
    finally:
        eng.set_tuning()


def test_graph_replay_matches_direct_launches(eng):
    """Steps replayed from the captured HIP graph give bit-identical states to
    direct launches, replay after replay, with eager work in between."""
    from particle_life import Engine, worlds

    print("\n[parity] case=graph_vs_direct", flush=True)
    w = worlds.uniform_world(116, 20_000, 6, rmax=0.02, dt=0.01)
    direct = Engine(0)
    direct.set_tuning(direct_launch=True)
    direct.load_world(w)
    eng.set_tuning()
    eng.load_world(w)
    for rep in range(4):
        eng.step(6)                      # three graph replays
        for _ in range(6):
            direct.step(1)
        a, b = eng.state(), direct.state()
        for k in ("pos", "vel", "ids"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"replay round {rep}: {k}")
        assert eng.pair_count() == direct.pair_count()
    direct.close()
