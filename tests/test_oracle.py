"""Pins the FP64 oracle (oracle/plife_oracle.c) without running the reference.

The reference has no tests or fixtures (nothing under src/test or among its
resources), and its Java cannot run here, so every check below comes from something the oracle's code
does not share: closed forms of the force law, hand-solved two-particle cases,
library routines (numpy's stable argsort / bincount), an independent numpy
all-pairs predicate, brute force, and invariants of the dynamics.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from particle_life import worlds


def _settings(**kw):
    base = dict(matrix=np.array([[1.0]]), rmax=0.05, friction=1.0, dt=0.02, beta=0.3)
    base.update(kw)
    return oracle.Settings(**base)


def _f_of(a, beta, dist):
    """Radial factor f recovered from plo_accelerate on rel = (dist, 0, 0)."""
    out = oracle.accelerate(a, beta, [dist, 0.0, 0.0])
    return out[0]  # acc_x = rel_x * f / dist = f for rel = (dist, 0, 0)


# ---------------------------------------------------------------- accelerator
@pytest.mark.parametrize("beta", [0.2, 0.25, 0.3, 0.4])
def test_accelerator_closed_forms(beta):
    a = 0.7
    assert abs(_f_of(a, beta, beta)) < 1e-15                          # f(beta) = 0
    assert abs(_f_of(a, beta, beta * (1 - 1e-12))) < 1e-11            # continuous from below
    assert _f_of(a, beta, (1 + beta) / 2) == pytest.approx(a, abs=1e-15)   # peak = a
    assert abs(_f_of(a, beta, 1.0)) < 1e-15                           # f(1) = 0
    assert _f_of(a, beta, 1e-9) == pytest.approx(-1.0, abs=1e-8)      # f(0+) = -1
    # repulsion does not depend on a; attraction is linear in a
    assert _f_of(-3.0, beta, beta / 2) == _f_of(5.0, beta, beta / 2) == pytest.approx(-0.5)
    d = (1 + beta) / 2 + 0.1
    assert _f_of(2 * a, beta, d) == pytest.approx(2 * _f_of(a, beta, d), rel=1e-15)


def test_accelerator_direction_and_hand_values():
    # beta = 0.3, rel = (0.3, 0.4, 0) -> dist = 0.5, f = 1 - |1.3 - 1|/0.7 = 4/7 (a = 1)
    out = oracle.accelerate(1.0, 0.3, [0.3, 0.4, 0.0])
    assert out == pytest.approx([0.3 * (4 / 7) / 0.5, 0.4 * (4 / 7) / 0.5, 0.0], rel=1e-14)
    # rel = (0, -0.1, 0): dist 0.1 < beta, f = 0.1/0.3 - 1 = -2/3, pushes away (+y)
    out = oracle.accelerate(-1.0, 0.3, [0.0, -0.1, 0.0])
    assert out == pytest.approx([0.0, 2 / 3, 0.0], rel=1e-14)


def test_friction_factor_closed_forms():
    assert oracle.friction_factor(0.5, 1 / 60) == pytest.approx(0.5, rel=1e-15)
    assert oracle.friction_factor(0.85, 0.0) == 1.0
    f = worlds.friction_from_half_life(0.04)
    assert oracle.friction_factor(f, 0.04) == pytest.approx(0.5, rel=1e-14)
    assert f == pytest.approx(0.7491535384383408, rel=1e-15)


# --------------------------------------------------------- binning (S2-S4, Q3)
def _keys_numpy(pos, rmax, dim):
    nx = math.floor(1.0 / rmax)
    c = np.trunc(pos[:, :dim] / rmax).astype(np.int64)
    c[c == nx] = nx - 1
    key = c[:, 0] + c[:, 1] * nx
    if dim == 3:
        key += c[:, 2] * nx * nx
    return key, nx


@pytest.mark.parametrize("rmax,dim", [(0.05, 2), (0.03, 2), (0.07, 2), (0.3, 2), (0.06, 3), (0.09, 3)])
def test_binning_matches_closed_form_and_stable_sort(rmax, dim):
    w = worlds.uniform_world(7, 3000, 5, rmax=rmax, dt=0.01, dim=dim)
    pos = w.pos.copy()
    pos[:7, 0] = [0.0, 1.0 - 2 ** -24, rmax, 2 * rmax, 0.5, 1.0, 0.25]   # cell edges, x = 1.0 (Q3)
    r = oracle.step(_settings(rmax=rmax, dim=dim, matrix=w.matrix), pos, w.vel, w.types, debug=True)
    key, nx = _keys_numpy(pos, rmax, dim)
    np.testing.assert_array_equal(r["keys"], key)
    np.testing.assert_array_equal(r["perm"], np.argsort(key, kind="stable"))
    ends = np.cumsum(np.bincount(key, minlength=nx ** dim))
    np.testing.assert_array_equal(r["cell_end"], ends)
    assert r["keys"][5] % nx == nx - 1     # x == 1.0 clamps to the last column


# ------------------------------------------------------------- neighbour sets
def _numpy_pairs(pos, rmax, wrap, dim):
    """Independent all-pairs predicate: counts of 0 < |d|^2 <= rmax^2 (FP64)."""
    d = pos[None, :, :] - pos[:, None, :]
    if wrap:
        d = np.where(d < -0.5, d + 1, np.where(d >= 0.5, d - 1, d))
    if dim == 2:
        d[..., 2] = 0.0
    d2 = d[..., 0] * d[..., 0] + (d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])
    inside = (d2 != 0) & (d2 <= rmax * rmax)
    np.fill_diagonal(inside, False)
    return inside.sum(axis=1)


@pytest.mark.parametrize("rmax,wrap,dim", [
    (0.05, True, 2), (0.03, True, 2), (0.07, True, 2), (0.11, False, 2),
    (0.25, True, 2), (0.30, True, 2), (0.08, True, 3), (0.12, False, 3),
])
def test_cell_list_neighbour_sets_equal_all_pairs(rmax, wrap, dim):
    w = worlds.uniform_world(11, 1500, 4, rmax=rmax, dt=0.01, dim=dim, wrap=wrap)
    S = _settings(rmax=rmax, wrap=wrap, dim=dim, matrix=w.matrix)
    r = oracle.step(S, w.pos, w.vel, w.types, debug=True)
    count, h = oracle.bruteforce_neighbors(S, w.pos)
    p = r["perm"]
    np.testing.assert_array_equal(r["nbr_count"], count[p])
    np.testing.assert_array_equal(r["nbr_hash"], h[p])
    np.testing.assert_array_equal(count, _numpy_pairs(w.pos, rmax, wrap, dim))


def test_repeated_stencil_cells_count_multiply():
    """Q1: with nx = 2 the 3x3 stencil visits a cell more than once and a pair is
    counted once per visit (Physics.java:377-395, 407-419)."""
    rmax = 0.4
    w = worlds.uniform_world(12, 300, 3, rmax=rmax, dt=0.01)
    S = _settings(rmax=rmax, matrix=w.matrix)
    r = oracle.step(S, w.pos, w.vel, w.types, debug=True)
    nx = 2
    pos = w.pos
    key, _ = _keys_numpy(pos, rmax, 2)
    cx = key % nx
    cy = key // nx
    o0x = np.floor(pos[:, 0] / rmax).astype(int)
    o0y = np.floor(pos[:, 1] / rmax).astype(int)
    d = pos[None, :, :] - pos[:, None, :]
    d = np.where(d < -0.5, d + 1, np.where(d >= 0.5, d - 1, d))
    d2 = d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]
    inside = (d2 != 0) & (d2 <= rmax * rmax)
    np.fill_diagonal(inside, False)
    expect = np.zeros(len(pos), np.int64)
    for i in range(len(pos)):
        sx = [((o0x[i] + o) % nx + nx) % nx for o in (-1, 0, 1)]
        sy = [((o0y[i] + o) % nx + nx) % nx for o in (-1, 0, 1)]
        js = np.nonzero(inside[i])[0]
        expect[i] = sum(sx.count(cx[j]) * sy.count(cy[j]) for j in js)
    np.testing.assert_array_equal(r["nbr_count"], expect[r["perm"]])


# ------------------------------------------------------- hand-solved two bodies
def test_two_particle_hand_solved():
    """p at (0.5, 0.5), q at rmax/2 to its right, v = 0, friction 0.9, dt = 1/64,
    force 2, M = [[a_pp, a_pq], [a_qp, a_qq]] = [[0, 0.75], [-0.5, 0]].
    |rel| = 1/2 >= beta = 0.3, so f = a (1 - |1.3 - 1|/0.7) = (4/7) a and
    dv = rel (f/|rel|) rmax force dt."""
    rmax, dt, force = 2.0 ** -5, 1 / 64, 2.0
    M = np.array([[0.0, 0.75], [-0.5, 0.0]])
    pos = np.array([[0.5, 0.5, 0.0], [0.5 + rmax / 2, 0.5, 0.0]])
    r = oracle.step(_settings(rmax=rmax, dt=dt, force=force, friction=0.9, matrix=M),
                    pos, np.zeros((2, 3)), np.array([0, 1], np.int32), debug=True)
    order = np.argsort(r["ids"])
    v = r["vel"][order]
    x = r["pos"][order]
    # p: rel = (+1/2, 0), f = 4/7 * 0.75 = 3/7, acc = rel * f / |rel| = (3/7, 0)
    vp = (3 / 7) * rmax * force * dt
    # q: rel = (-1/2, 0), f = 4/7 * (-0.5) = -2/7, acc = rel * f / |rel| = (+2/7, 0)
    vq = (2 / 7) * rmax * force * dt
    assert v[0] == pytest.approx([vp, 0, 0], abs=1e-18)
    assert v[1] == pytest.approx([vq, 0, 0], abs=1e-18)
    assert x[0] == pytest.approx([0.5 + vp * dt, 0.5, 0], abs=1e-17)
    assert x[1] == pytest.approx([0.5 + rmax / 2 + vq * dt, 0.5, 0], abs=1e-17)
    assert r["pair_count"] == 2


def test_two_particle_repulsion_across_seam():
    """p at x = 1 - rmax/20, q at x = rmax/20 (distance rmax/10 through the seam):
    f = 0.1/0.3 - 1 = -2/3, each is pushed away from the other."""
    rmax, dt = 2.0 ** -4, 1 / 32
    pos = np.array([[1 - rmax / 20, 0.25, 0.0], [rmax / 20, 0.25, 0.0]])
    r = oracle.step(_settings(rmax=rmax, dt=dt, friction=1.0, matrix=np.array([[1.0]])),
                    pos, np.zeros((2, 3)), np.array([0, 0], np.int32), debug=True)
    order = np.argsort(r["ids"])
    v = r["vel"][order]
    mag = (2 / 3) * rmax * dt
    assert v[0] == pytest.approx([-mag, 0, 0], rel=1e-12)   # pushed to -x
    assert v[1] == pytest.approx([+mag, 0, 0], rel=1e-12)   # pushed to +x


# -------------------------------------------------------------- integrate/wrap
def test_wrap_and_clamp_special_cases():
    dt = 2.0 ** -6
    pos = np.array([[0.0, 0.5, 0], [0.5, 1 - 2.0 ** -20, 0], [0.999, 0.25, 0]])
    vel = np.array([[-(2.0 ** -54), 0, 0], [0, 2.0 ** -14, 0], [1 / dt, 0, 0]])
    far = dict(rmax=0.01, dt=dt, friction=1.0)
    r = oracle.step(_settings(**far), pos, vel, np.zeros(3, np.int32))
    x = r["pos"][np.argsort(r["ids"])]
    assert x[0, 0] == 1.0                     # -2^-60 + 1 rounds to exactly 1.0 (Q3)
    assert x[1, 1] == 0.0                     # (1 - 2^-20) + 2^-20 = 1.0 wraps to 0
    assert x[2, 0] == pytest.approx(0.999)    # 0.999 + 1 -> wraps by repeated -1
    r = oracle.step(_settings(wrap=False, **far), pos, vel, np.zeros(3, np.int32))
    x = r["pos"][np.argsort(r["ids"])]
    assert x[0, 0] == 0.0 and x[2, 0] == 1.0  # clamped to [0, 1]
    # a particle sitting at exactly 1.0 is keyed to the last cell and then wraps to 0
    pos1 = np.array([[1.0, 0.5, 0.0]])
    r = oracle.step(_settings(**far), pos1, np.zeros((1, 3)), np.zeros(1, np.int32), debug=True)
    assert r["keys"][0] == 99 + 50 * 100 and r["pos"][0, 0] == 0.0


# ------------------------------------------------------------------ invariants
def test_momentum_conserved_with_symmetric_matrix():
    w = worlds.uniform_world(13, 4000, 5, rmax=0.03, dt=0.02)
    m = w.matrix + w.matrix.T
    vel = np.zeros((w.n, 3))
    vel[:, :2] = worlds.uniform24(99, 1, 2 * w.n).reshape(-1, 2) - 0.5
    vel[:, :2] -= vel[:, :2].mean(axis=0)
    r = oracle.step(_settings(rmax=0.03, friction=1.0, matrix=m), w.pos, vel, w.types)
    total = r["vel"].sum(axis=0)
    assert np.all(np.abs(total - vel.sum(axis=0)) < 1e-12 * np.abs(r["vel"]).sum())


def test_translation_by_whole_cells_gives_identical_results():
    rmax = 2.0 ** -5
    w = worlds.uniform_world(14, 3000, 4, rmax=rmax, dt=0.01)
    S = _settings(rmax=rmax, matrix=w.matrix, friction=0.8)
    r0 = oracle.step(S, w.pos, w.vel, w.types, debug=True)
    shifted = w.pos.copy()
    shifted[:, :2] = (shifted[:, :2] + np.array([5 * rmax, 11 * rmax])) % 1.0
    r1 = oracle.step(S, shifted, w.vel, w.types, debug=True)
    o0, o1 = np.argsort(r0["ids"]), np.argsort(r1["ids"])
    np.testing.assert_array_equal(r0["nbr_count"][o0], r1["nbr_count"][o1])
    np.testing.assert_array_equal(r0["vel"][o0], r1["vel"][o1])


def test_thread_count_and_input_order_independence():
    w = worlds.config1(5000)
    S = _settings(rmax=w.rmax, matrix=w.matrix, friction=w.friction, dt=w.dt)
    a = oracle.step(S, w.pos, w.vel, w.types, nthreads=1)
    b = oracle.step(S, w.pos, w.vel, w.types, nthreads=7)
    for k in ("pos", "vel", "ids"):
        np.testing.assert_array_equal(a[k], b[k])
    perm = np.random.default_rng(1).permutation(w.n)
    c = oracle.step(S, w.pos[perm], w.vel[perm], w.types[perm], ids=perm.astype(np.int32))
    oa, oc = np.argsort(a["ids"]), np.argsort(c["ids"])
    np.testing.assert_allclose(c["vel"][oc], a["vel"][oa], rtol=1e-12, atol=1e-18)


def test_3d_with_flat_z_equals_2d():
    w = worlds.uniform_world(15, 3000, 5, rmax=0.05, dt=0.02)
    S2 = _settings(rmax=0.05, matrix=w.matrix, friction=0.9)
    S3 = _settings(rmax=0.05, matrix=w.matrix, friction=0.9, dim=3)
    a = oracle.step(S2, w.pos, w.vel, w.types, debug=True)
    b = oracle.step(S3, w.pos, w.vel, w.types, debug=True)
    oa, ob = np.argsort(a["ids"]), np.argsort(b["ids"])
    np.testing.assert_array_equal(a["vel"][oa], b["vel"][ob])
    np.testing.assert_array_equal(a["nbr_hash"][oa], b["nbr_hash"][ob])


def test_bad_arguments_are_rejected():
    with pytest.raises(ValueError):
        oracle.step(_settings(rmax=0.0), np.zeros((1, 3)), np.zeros((1, 3)), np.zeros(1, np.int32))
    with pytest.raises(ValueError):
        oracle.step(_settings(), np.zeros((1, 3)), np.zeros((1, 3)), np.array([3], np.int32))
