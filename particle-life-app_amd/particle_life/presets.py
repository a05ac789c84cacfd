"""The app's position, type and matrix presets.

Names and order follow the app's selection lists
(app/PositionSetterProvider.java:17-119, app/TypeSetterProvider.java:12-45,
app/MatrixGeneratorProvider.java:14-79).  Each setter works two ways:

* per particle on the host, through the reference's PositionSetter.set /
  TypeSetter.getType interface (draws from a seeded numpy Generator, as
  ``DefaultPositionSetter`` does in this package) -- what
  ``GpuPhysics.setParticleCount`` uses for new particles;
* for all particles at once on the device: ``GpuPhysics.setPositions()`` /
  ``setTypes()`` call ``plife_set_positions`` / ``plife_set_types`` when the
  active setter is one of these presets (``Engine.set_positions``,
  ``Batch.set_positions`` and friends take the preset directly).

Matrix presets are host-only (``plife_generate_matrix``).
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional, Union

import numpy as np

from . import _native
from ._native import check
from .physics import DefaultMatrix, MatrixGenerator, PositionSetter, TypeSetter

POSITION_PRESETS = ("centered", "uniform", "uniform circle", "centered circle", "ring",
                    "rainbow ring", "color battle", "color wheel", "line", "spiral",
                    "rainbow spiral")
TYPE_PRESETS = ("random", "randomize 10%", "slices", "onion", "rotate", "flip", "more of first",
                "kill still")
MATRIX_PRESETS = ("random", "symmetry", "chains", "chains 2", "chains 3", "snakes", "zero")

_F03, _F05, _F01, _F02, _F015 = (float(np.float32(v)) for v in (0.3, 0.5, 0.1, 0.2, 0.15))


def preset_index(names: tuple, preset: Union[int, str]) -> int:
    """Index of `preset` (a name or an index) in `names`."""
    if isinstance(preset, str):
        if preset not in names:
            raise ValueError(f"unknown preset {preset!r}; one of {names}")
        return names.index(preset)
    i = int(preset)
    if not 0 <= i < len(names):
        raise ValueError(f"preset index {i} out of range [0, {len(names)})")
    return i


def generate_matrix(preset: Union[int, str], size: int, seed: int) -> np.ndarray:
    """plife_generate_matrix: the (size, size) matrix of a matrix preset."""
    out = np.empty((int(size), int(size)), dtype=np.float64)
    check(_native.load().plife_generate_matrix(preset_index(MATRIX_PRESETS, preset), int(size),
                                              ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                              out.ctypes.data))
    return out


class _Seeded:
    """Seed bookkeeping: every device call draws from a fresh seed, as the
    reference's setters draw fresh random numbers on every call."""

    def __init__(self, seed: int):
        self.seed = int(seed) & (2**64 - 1)
        self.calls = 0

    def next_seed(self) -> int:
        s = (self.seed + self.calls * 0x9E3779B97F4A7C15) & (2**64 - 1)
        self.calls += 1
        return s


class PresetPositionSetter(PositionSetter, _Seeded):
    """One of the app's position presets (PositionSetterProvider order)."""

    def __init__(self, preset: Union[int, str] = "uniform", seed: int = 0,
                 rng: Optional[np.random.Generator] = None):
        _Seeded.__init__(self, seed)
        self.index = preset_index(POSITION_PRESETS, preset)
        self.name = POSITION_PRESETS[self.index]
        self._rng = rng or np.random.default_rng(self.seed)

    def set(self, position: np.ndarray, type_: int, nTypes: int) -> None:
        r = self._rng
        u, g = r.random, r.standard_normal
        name = self.name
        if name == "uniform":           # DefaultPositionSetter
            position[0], position[1], position[2] = u(), u(), 0.0
            return
        if name == "centered":
            x, y = g() * _F03, g() * _F03
        elif name in ("uniform circle", "centered circle"):
            angle = u() * 2 * math.pi
            radius = 0.5 * math.sqrt(u()) if name == "uniform circle" else _F05 * u()
            x, y = math.cos(angle) * radius, math.sin(angle) * radius
        elif name == "ring":
            angle = u() * 2 * math.pi
            radius = 0.7 + 0.02 * g()
            x, y = math.cos(angle) * radius, math.sin(angle) * radius
        elif name == "rainbow ring":
            angle = (0.3 * g() + type_) / nTypes * 2 * math.pi
            radius = 0.7 + 0.02 * g()
            x, y = math.cos(angle) * radius, math.sin(angle) * radius
        elif name == "color battle":
            ca = type_ / nTypes * 2 * math.pi
            angle = u() * 2 * math.pi
            radius = u() * _F01
            x = _F05 * math.cos(ca) + math.cos(angle) * radius
            y = _F05 * math.sin(ca) + math.sin(angle) * radius
        elif name == "color wheel":
            ca = type_ / nTypes * 2 * math.pi
            x = _F03 * math.cos(ca) + g() * _F02
            y = _F03 * math.sin(ca) + g() * _F02
        elif name == "line":
            x, y = 2 * u() - 1, (2 * u() - 1) * _F015
        else:                           # spiral, rainbow spiral
            if name == "spiral":
                f = u()
            else:
                f = (type_ + 1) / (nTypes + 2) + 0.3 / nTypes * g()
                f = min(max(f, 0.0), 1.0)
            angle = 2.0 * 2 * math.pi * f
            spread = 0.5 * min(f, 0.2)
            radius = 0.9 * f + spread * g() * spread
            x, y = radius * math.cos(angle), radius * math.sin(angle)
        position[0], position[1] = x * 0.5 + 0.5, y * 0.5 + 0.5


def _map_type(v: float, n: int) -> int:
    return max(0, min(n - 1, int(math.floor(v * n))))


class PresetTypeSetter(TypeSetter, _Seeded):
    """One of the app's type presets (TypeSetterProvider order)."""

    def __init__(self, preset: Union[int, str] = "random", seed: int = 0,
                 rng: Optional[np.random.Generator] = None):
        _Seeded.__init__(self, seed)
        self.index = preset_index(TYPE_PRESETS, preset)
        self.name = TYPE_PRESETS[self.index]
        self._rng = rng or np.random.default_rng(self.seed)

    def getType(self, position, velocity, type_: int, nTypes: int) -> int:
        u = self._rng.random
        name = self.name
        if name == "random":
            return int(math.floor(u() * nTypes))
        if name == "randomize 10%":
            return _map_type(u(), nTypes) if u() < 0.1 else type_
        if name == "slices":
            return _map_type(position[0], nTypes)
        if name == "onion":
            dx, dy, dz = position[0] - 0.5, position[1] - 0.5, position[2]
            return _map_type(math.sqrt(dx * dx + (dy * dy + dz * dz)) * 2, nTypes)
        if name == "rotate":
            return (type_ + 1) % nTypes
        if name == "flip":
            return nTypes - 1 - type_
        if name == "more of first":
            return _map_type(u() * u(), nTypes)
        vx, vy, vz = velocity[0], velocity[1], velocity[2]           # kill still
        return nTypes - 1 if math.sqrt(vx * vx + (vy * vy + vz * vz)) < 0.01 else type_


class PresetMatrixGenerator(MatrixGenerator, _Seeded):
    """One of the app's matrix presets (MatrixGeneratorProvider order)."""

    def __init__(self, preset: Union[int, str] = "random", seed: int = 0):
        _Seeded.__init__(self, seed)
        self.index = preset_index(MATRIX_PRESETS, preset)
        self.name = MATRIX_PRESETS[self.index]

    def makeMatrix(self, size: int) -> DefaultMatrix:
        return DefaultMatrix(int(size), generate_matrix(self.index, int(size), self.next_seed()))
