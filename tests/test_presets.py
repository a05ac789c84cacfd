"""The presets' oracle pinned by what the shapes fix, and the host-only
plife_generate_matrix against it (CPU; the device kernels are checked in
test_gpu_presets.py)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import presets as op

N, T, SEED = 20_000, 6, 0x5EED0077
ALL_TYPES = np.arange(N, dtype=np.int32) % T


def xy(name, types=ALL_TYPES, ntypes=T, seed=SEED):
    return op.shape_positions(op.POSITION_PRESETS.index(name), types, ntypes, seed)


def polar(x, y):
    dx, dy = x - 0.5, y - 0.5
    return np.hypot(dx, dy), np.mod(np.arctan2(dy, dx), 2 * np.pi)


def test_uniform_preset_is_the_host_generator():
    from particle_life import worlds
    w = worlds.uniform_world(SEED, N, T, rmax=0.02, dt=0.01)
    px, py, pz = op.set_positions(op.POSITION_PRESETS.index("uniform"), ALL_TYPES, T, SEED, 2, True)
    np.testing.assert_array_equal(px.astype(np.float64) * 2.0 ** -32, w.pos[:, 0])
    np.testing.assert_array_equal(py.astype(np.float64) * 2.0 ** -32, w.pos[:, 1])
    assert not pz.any()
    w3 = worlds.uniform_world(SEED, N, T, rmax=0.02, dt=0.01, dim=3)
    _, _, pz3 = op.set_positions(op.POSITION_PRESETS.index("uniform"), ALL_TYPES, T, SEED, 3, True)
    np.testing.assert_array_equal(pz3.astype(np.float64) * 2.0 ** -32, w3.pos[:, 2])


def test_draws_do_not_depend_on_the_count():
    for name in op.POSITION_PRESETS:
        a = xy(name)
        b = xy(name, types=ALL_TYPES[:100])
        np.testing.assert_array_equal(a[0][:100], b[0])
        np.testing.assert_array_equal(a[1][:100], b[1])


def test_gaussians_are_standard():
    g = op.gauss(SEED, 40, 200_000)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1) < 0.01
    assert abs(np.mean(np.abs(g) < 1) - 0.6827) < 0.005


def test_centered():
    x, y = xy("centered")
    for v in (x, y):                   # N(0.5, (0.3f * 0.5)^2)
        assert abs(v.mean() - 0.5) < 0.005 and abs(v.std() - 0.15) < 0.003


def test_circles():
    r, _ = polar(*xy("uniform circle"))
    assert r.max() <= 0.25 + 1e-12
    assert abs(np.mean(r < 0.125) - 0.25) < 0.01          # uniform in area
    r, _ = polar(*xy("centered circle"))
    assert r.max() <= 0.25 + 1e-12
    assert abs(np.mean(r < 0.125) - 0.5) < 0.01           # uniform in radius


def test_rings():
    for name in ("ring", "rainbow ring"):
        r, _ = polar(*xy(name))
        assert abs(r.mean() - 0.35) < 0.001 and abs(r.std() - 0.01) < 0.001
    _, a = polar(*xy("rainbow ring"))
    for t in range(T):                 # circular mean angle of type t: t / T * 2 pi
        sel = ALL_TYPES == t
        m = math.atan2(np.sin(a[sel]).mean(), np.cos(a[sel]).mean()) % (2 * np.pi)
        d = (m - t / T * 2 * np.pi + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) < 0.02


def test_color_battle_and_wheel():
    x, y = xy("color battle")
    for t in range(T):
        ca = t / T * 2 * np.pi
        sel = ALL_TYPES == t
        cx, cy = 0.5 + 0.25 * math.cos(ca), 0.5 + 0.25 * math.sin(ca)
        assert np.hypot(x[sel] - cx, y[sel] - cy).max() <= 0.05 + 1e-9
    x, y = xy("color wheel")
    for t in range(T):
        ca = t / T * 2 * np.pi
        sel = ALL_TYPES == t
        assert abs(x[sel].mean() - (0.5 + op.F03 * 0.5 * math.cos(ca))) < 0.01
        assert abs(y[sel].mean() - (0.5 + op.F03 * 0.5 * math.sin(ca))) < 0.01
        assert abs(x[sel].std() - op.F02 * 0.5) < 0.005


def test_line():
    x, y = xy("line")
    assert np.abs(y - 0.5).max() <= op.F015 * 0.5 + 1e-12
    assert x.min() >= 0 and x.max() < 1 and abs(x.mean() - 0.5) < 0.01


def test_spirals():
    r, a = polar(*xy("spiral"))
    # Archimedean spiral of two turns: r = 0.45 * theta / (4 pi), theta = a + 2 pi k
    resid = np.min([np.abs(r - 0.45 * (a + 2 * np.pi * k) / (4 * np.pi)) for k in (0, 1)], axis=0)
    assert np.mean(resid < 0.03) > 0.99
    r, _ = polar(*xy("rainbow spiral"))
    for t in range(T):
        f = (t + 1) / (T + 2)
        assert abs(np.median(r[ALL_TYPES == t]) - 0.45 * f) < 0.02


def _fixed_pos(name, dim=2, wrap=True):
    return op.set_positions(op.POSITION_PRESETS.index(name), ALL_TYPES, T, SEED, dim, wrap)


def test_type_presets_closed_forms():
    px, py, pz = _fixed_pos("uniform")
    vel = np.zeros((N, 3), np.float32)
    f = lambda name, types=ALL_TYPES, v=vel: op.set_types(op.TYPE_PRESETS.index(name), px, py, pz,
                                                          v, types, T, SEED, 2)
    np.testing.assert_array_equal(f("rotate"), (ALL_TYPES + 1) % T)
    np.testing.assert_array_equal(f("flip"), T - 1 - ALL_TYPES)
    from particle_life import worlds
    np.testing.assert_array_equal(f("random"), worlds.uniform_world(SEED, N, T, 0.02, 0.01).types)
    x = px.astype(np.float64) * 2.0 ** -32
    s = f("slices")
    order = np.argsort(x, kind="stable")
    assert np.all(np.diff(s[order]) >= 0)                 # type grows with x ...
    for t in range(T):                                    # ... in bands of width 1 / T
        assert x[s == t].min() >= t / T - 1e-12 and x[s == t].max() < (t + 1) / T
    r = np.hypot(x - 0.5, py.astype(np.float64) * 2.0 ** -32 - 0.5)
    o = f("onion")
    assert np.all(np.diff(o[np.argsort(r, kind="stable")]) >= 0)
    changed = f("randomize 10%") != ALL_TYPES
    assert abs(changed.mean() - 0.1 * (1 - 1 / T)) < 0.006
    first = f("more of first")
    assert abs(np.mean(first == 0) - (1 + math.log(T)) / T) < 0.01     # P(u1 u2 < 1/T)
    v = np.zeros((N, 3), np.float32)
    v[::2, 0] = 0.02
    v[1::4, 1] = 0.009
    k = f("kill still", v=v)
    np.testing.assert_array_equal(k[::2], ALL_TYPES[::2])
    assert np.all(k[1::2] == T - 1)


@pytest.mark.parametrize("size", [1, 2, 3, 7, 12])
def test_matrix_presets(size):
    from particle_life import presets, worlds
    for i, name in enumerate(op.MATRIX_PRESETS):
        m = presets.generate_matrix(name, size, SEED)
        np.testing.assert_array_equal(m, op.matrix(i, size, SEED), err_msg=name)
    rnd = op.matrix(0, size, SEED)
    np.testing.assert_array_equal(rnd, worlds.uniform_world(SEED, 1, size, 0.02, 0.01).matrix)
    sym = op.matrix(1, size, SEED)
    np.testing.assert_array_equal(sym, sym.T)
    np.testing.assert_array_equal(np.tril(sym), np.tril(rnd))
    ch = op.matrix(2, size, SEED)
    for i in range(size):
        for j in range(size):
            near = (i - j) % size in (0, 1, size - 1)
            assert ch[i, j] == (1.0 if near else -1.0)
    assert not op.matrix(6, size, SEED).any()


def test_generate_matrix_rejects_bad_arguments():
    from particle_life import presets
    from particle_life._native import PlifeError
    with pytest.raises(PlifeError):
        presets.generate_matrix(0, 0, 1)
    with pytest.raises(PlifeError):
        presets.generate_matrix(0, 257, 1)
    with pytest.raises(ValueError):
        presets.generate_matrix("no such preset", 3, 1)


def test_host_preset_setters_follow_the_shapes():
    from particle_life import presets
    ps = presets.PresetPositionSetter("ring", seed=3)
    pts = np.zeros((2000, 3))
    for i in range(2000):
        ps.set(pts[i], i % T, T)
    r = np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5)
    assert abs(r.mean() - 0.35) < 0.002
    ts = presets.PresetTypeSetter("flip")
    assert ts.getType(np.zeros(3), np.zeros(3), 2, T) == T - 3
