"""The app's position, type and matrix presets with this build's seeded draws,
in numpy -- TEST INFRASTRUCTURE ONLY (the checker of plife_set_positions,
plife_set_types and plife_generate_matrix; nothing in the product imports it).

Follows app/PositionSetterProvider.java:17-119, app/TypeSetterProvider.java:12-45,
app/MatrixGeneratorProvider.java:14-79, Physics.setPosition / ensurePosition
(backend/Physics.java:297-303, 499-505) and DefaultMatrix.randomize
(backend/DefaultMatrix.java:24-30), with the reference's unseeded
java.util.Random replaced by counter-based SplitMix64 streams of a seed:

* draw k of stream s is mix(mix(seed + s * C) + (k + 1) * G) (Vigna's
  SplitMix64, C = 0xD1B54A32D192ED03, G = 0x9E3779B97F4A7C15);
* preset streams are s = 40..45; the uniform generator's are s = 1..5
  (x, y, z, type, matrix), on the 2^-24 grid (top 24 bits);
* other uniforms take the top 53 bits; Gaussians are Box-Muller,
  sqrt(-2 ln u1) cos(2 pi u2) with u1 = (top53 + 1) 2^-53 in (0, 1].

Parity status: every preset is pinned by properties the mathematics fixes
(tests/test_presets.py): shapes (radii, bands, per-type centres, angles),
closed forms of the integer presets (rotate, flip, slices, the chain
matrices), the uniform preset and random matrix against the independent host
generator (worlds.uniform_world), draw-count invariance (a prefix of n draws
does not depend on n).  The device uses the same formulas with its own libm;
sin, cos, log and sqrt may differ by an ulp, so device positions are compared
within 2 units of 2^-32.  The reference's own draws cannot be reproduced
(unseeded), so the distributions, not the values, are the reference's.
"""
from __future__ import annotations

import numpy as np

G = np.uint64(0x9E3779B97F4A7C15)
C = np.uint64(0xD1B54A32D192ED03)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
F03, F05, F01, F02, F015 = (float(np.float32(v)) for v in (0.3, 0.5, 0.1, 0.2, 0.15))
POSITION_PRESETS = ("centered", "uniform", "uniform circle", "centered circle", "ring",
                    "rainbow ring", "color battle", "color wheel", "line", "spiral",
                    "rainbow spiral")
TYPE_PRESETS = ("random", "randomize 10%", "slices", "onion", "rotate", "flip", "more of first",
                "kill still")
MATRIX_PRESETS = ("random", "symmetry", "chains", "chains 2", "chains 3", "snakes", "zero")


def _mix(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def draws(seed: int, s: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = _mix(np.array([np.uint64(seed & (2**64 - 1)) + np.uint64(s) * C], np.uint64))[0]
        return _mix(base + np.arange(1, n + 1, dtype=np.uint64) * G)


def u53(seed, s, n):
    return (draws(seed, s, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def u24(seed, s, n):
    return (draws(seed, s, n) >> np.uint64(40)).astype(np.float64) * 2.0 ** -24


def gauss(seed, s, n):
    u1 = ((draws(seed, s, n) >> np.uint64(11)).astype(np.float64) + 1.0) * 2.0 ** -53
    u2 = (draws(seed, s + 1, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def fixed(x: np.ndarray, wrap: bool) -> np.ndarray:
    """ensurePosition, then rounding to the 2^-32 grid (u32; periodic 1.0 -> 0)."""
    x = np.asarray(x, np.float64)
    x = np.where(np.isnan(x), 0.0, x)
    if wrap:
        x = x - np.floor(x)
        return np.rint(x * 2.0 ** 32).astype(np.uint64).astype(np.uint32)   # 2^32 wraps to 0
    x = np.clip(x, 0.0, 1.0)
    return np.minimum(np.rint(x * 2.0 ** 32), 2.0 ** 32 - 1).astype(np.uint64).astype(np.uint32)


def shape_positions(preset: int, types: np.ndarray, ntypes: int, seed: int):
    """(x, y) of a position preset in the unit box, before ensurePosition."""
    name = POSITION_PRESETS[preset]
    n = types.shape[0]
    t = types.astype(np.float64)
    T = float(ntypes)
    if name == "uniform":
        return u24(seed, 1, n), u24(seed, 2, n)
    if name == "centered":
        x, y = gauss(seed, 40, n) * F03, gauss(seed, 42, n) * F03
    elif name in ("uniform circle", "centered circle"):
        angle = u53(seed, 40, n) * 2 * np.pi
        r = 0.5 * np.sqrt(u53(seed, 41, n)) if name == "uniform circle" else F05 * u53(seed, 41, n)
        x, y = np.cos(angle) * r, np.sin(angle) * r
    elif name == "ring":
        angle = u53(seed, 40, n) * 2 * np.pi
        r = 0.7 + 0.02 * gauss(seed, 42, n)
        x, y = np.cos(angle) * r, np.sin(angle) * r
    elif name == "rainbow ring":
        angle = (0.3 * gauss(seed, 40, n) + t) / T * 2 * np.pi
        r = 0.7 + 0.02 * gauss(seed, 42, n)
        x, y = np.cos(angle) * r, np.sin(angle) * r
    elif name == "color battle":
        ca = t / T * 2 * np.pi
        angle = u53(seed, 40, n) * 2 * np.pi
        r = u53(seed, 41, n) * F01
        x = F05 * np.cos(ca) + np.cos(angle) * r
        y = F05 * np.sin(ca) + np.sin(angle) * r
    elif name == "color wheel":
        ca = t / T * 2 * np.pi
        x = F03 * np.cos(ca) + gauss(seed, 40, n) * F02
        y = F03 * np.sin(ca) + gauss(seed, 42, n) * F02
    elif name == "line":
        x = 2 * u53(seed, 40, n) - 1
        y = (2 * u53(seed, 41, n) - 1) * F015
    else:
        if name == "spiral":
            f = u53(seed, 40, n)
        else:
            f = (t + 1) / float(ntypes + 2) + (0.3 / ntypes) * gauss(seed, 40, n)
            f = np.clip(f, 0.0, 1.0)
        angle = 2.0 * 2 * np.pi * f
        spread = 0.5 * np.minimum(f, 0.2)
        r = 0.9 * f + spread * gauss(seed, 42, n) * spread
        x, y = r * np.cos(angle), r * np.sin(angle)
    return x * 0.5 + 0.5, y * 0.5 + 0.5


def set_positions(preset: int, types, ntypes: int, seed: int, dim: int, wrap: bool):
    """Fixed-point (px, py, pz) after plife_set_positions on particles of `types`."""
    types = np.asarray(types)
    x, y = shape_positions(preset, types, ntypes, seed)
    n = types.shape[0]
    z = ((draws(seed, 3, n) >> np.uint64(40)) << np.uint64(8)).astype(np.uint32) if dim == 3 \
        else np.zeros(n, np.uint32)
    return fixed(x, wrap), fixed(y, wrap), z


def _map_type(v, nt):
    return np.clip(np.floor(v * nt), 0, nt - 1).astype(np.int32)


def set_types(preset: int, px, py, pz, vel, types, ntypes: int, seed: int, dim: int):
    """Types after plife_set_types; positions as u32 fixed point, vel (n, 3) f32."""
    name = TYPE_PRESETS[preset]
    types = np.asarray(types, np.int32)
    n = types.shape[0]
    nt = int(ntypes)
    if name == "random":
        m = draws(seed, 4, n) >> np.uint64(40)
        return ((m * np.uint64(nt)) >> np.uint64(24)).astype(np.int32)
    if name == "randomize 10%":
        return np.where(u53(seed, 44, n) < 0.1, _map_type(u53(seed, 45, n), nt), types).astype(np.int32)
    x = np.asarray(px).astype(np.float64) * 2.0 ** -32
    y = np.asarray(py).astype(np.float64) * 2.0 ** -32
    z = np.asarray(pz).astype(np.float64) * 2.0 ** -32 if dim == 3 else np.zeros(n)
    if name == "slices":
        return _map_type(x, nt)
    if name == "onion":
        dx, dy = x - 0.5, y - 0.5
        return _map_type(np.sqrt(dx * dx + (dy * dy + z * z)) * 2, nt)
    if name == "rotate":
        return ((types + 1) % nt).astype(np.int32)
    if name == "flip":
        return (nt - 1 - types).astype(np.int32)
    if name == "more of first":
        return _map_type(u53(seed, 44, n) * u53(seed, 45, n), nt)
    v = np.asarray(vel, np.float32).astype(np.float64)
    vz = v[:, 2] if dim == 3 else np.zeros(n)
    speed = np.sqrt(v[:, 0] * v[:, 0] + (v[:, 1] * v[:, 1] + vz * vz))
    return np.where(speed < 0.01, nt - 1, types).astype(np.int32)


def matrix(preset: int, size: int, seed: int) -> np.ndarray:
    name = MATRIX_PRESETS[preset]
    n = int(size)
    if name in ("random", "symmetry"):
        m = (2.0 * u24(seed, 5, n * n) - 1.0).reshape(n, n)
        if name == "symmetry":
            for i in range(n):
                for j in range(i, n):
                    m[i, j] = m[j, i]
        return m
    m = np.zeros((n, n))
    if name == "snakes":
        for i in range(n):
            m[i, i] = 1.0
            m[i, (i + 1) % n] = 0.2
        return m
    if name == "zero":
        return m
    for i in range(n):
        for j in range(n):
            near = j == (i + 1) % n or j == (i + n - 1) % n
            if name == "chains":
                m[i, j] = 1.0 if (j == i or near) else -1.0
            else:
                m[i, j] = 1.0 if j == i else (0.2 if near else (-1.0 if name == "chains 2" else 0.0))
    return m
