"""FP64 CPU oracle for particle-life's physics step -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this package, and only as the checker (or the timed CPU baseline); the
product path (the HIP library and its host mirror) never touches it.

The C restatement lives in ``plife_oracle.c`` (every rule cites the reference
file:line it follows); this module is a thin ctypes wrapper around
``liboracle.so`` built by ``oracle/Makefile``.

Parity status of each oracle function (how it is pinned, see tests/):
  * plo_accelerate      -- closed forms f(beta)=0, f((1+beta)/2)=a, f(1)=0,
                           f(0+)=-1, hand-solved two-particle cases.
  * cell keys / ranges  -- closed form trunc(x/cs) with clamp, numpy's stable
                           argsort for the permutation, brute-force enumeration.
  * neighbour sets      -- O(n^2) all-pairs predicate (plo_bruteforce_neighbors)
                           plus an independent numpy all-pairs restatement.
  * velocity / position -- two-particle hand-solved cases, momentum
                           conservation with symmetric matrix and no friction,
                           translation and input-order invariance.
  * JOML rounding order (lengthSquared, div, mulAdd) and Java Math.pow's last
    ulp -- parity unpinned (no JDK or JOML here; affects only the last FP64 ulp).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class _Settings(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("wrap", ctypes.c_int),
        ("rmax", ctypes.c_double),
        ("friction", ctypes.c_double),
        ("force", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("ntypes", ctypes.c_int),
        ("matrix", ctypes.POINTER(ctypes.c_double)),
    ]


class _Debug(ctypes.Structure):
    _fields_ = [
        ("keys", ctypes.POINTER(ctypes.c_int32)),
        ("cell_end", ctypes.POINTER(ctypes.c_int32)),
        ("perm", ctypes.POINTER(ctypes.c_int32)),
        ("nbr_count", ctypes.POINTER(ctypes.c_int32)),
        ("nbr_hash", ctypes.POINTER(ctypes.c_uint64)),
        ("pair_count", ctypes.c_int64),
        ("seconds", ctypes.c_double),
    ]


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.plo_step.restype = ctypes.c_int
        L.plo_step.argtypes = [
            ctypes.POINTER(_Settings), ctypes.c_int,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int, ctypes.POINTER(_Debug),
        ]
        L.plo_bruteforce_neighbors.restype = ctypes.c_int
        L.plo_bruteforce_neighbors.argtypes = [
            ctypes.POINTER(_Settings), ctypes.c_int, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
        ]
        L.plo_accelerate.restype = None
        L.plo_accelerate.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
        L.plo_id_hash.restype = ctypes.c_uint64
        L.plo_id_hash.argtypes = [ctypes.c_uint32]
        L.plo_friction_factor.restype = ctypes.c_double
        L.plo_friction_factor.argtypes = [ctypes.c_double, ctypes.c_double]
        L.plo_cells_per_axis.restype = ctypes.c_int
        L.plo_cells_per_axis.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


@dataclass
class Settings:
    """Mirror of PhysicsSettings (backend/PhysicsSettings.java:8-38) plus the
    accelerator's beta (app/Main.java:276) and the world dimension."""

    matrix: np.ndarray
    rmax: float = 0.02
    friction: float = 0.85
    force: float = 1.0
    dt: float = 0.02
    beta: float = 0.3
    wrap: bool = True
    dim: int = 2
    _keep: list = field(default_factory=list, repr=False)

    def _c(self) -> _Settings:
        m = np.ascontiguousarray(self.matrix, dtype=np.float64)
        self._keep = [m]
        s = _Settings()
        s.dim = int(self.dim)
        s.wrap = 1 if self.wrap else 0
        s.rmax = float(self.rmax)
        s.friction = float(self.friction)
        s.force = float(self.force)
        s.dt = float(self.dt)
        s.beta = float(self.beta)
        s.ntypes = int(m.shape[0])
        s.matrix = m.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return s


def cells_per_axis(rmax: float) -> int:
    return int(lib().plo_cells_per_axis(float(rmax)))


def ncells(settings: Settings) -> int:
    nx = cells_per_axis(settings.rmax)
    return nx ** settings.dim


def step(settings: Settings, pos, vel, types, ids=None, nthreads: int = 1, debug: bool = False,
         hashes: bool = True):
    """One Physics.update().  Returns a dict with the new state in the new
    (cell-sorted) array order and, with debug=True, keys / cell_end / perm /
    nbr_count / nbr_hash / pair_count and seconds, the wall time of the step
    proper inside plo_step (binning and the two parallel phases, without the
    copies in and out of its particle records); hashes=False leaves out
    nbr_hash and its per-pair work, for timing."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3).copy()
    vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3).copy()
    types = np.ascontiguousarray(types, dtype=np.int32).copy()
    n = types.shape[0]
    ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, dtype=np.int32).copy()
    cs = settings._c()
    dbg = None
    out = {}
    if debug:
        nc = ncells(settings)
        out["keys"] = np.zeros(n, np.int32)
        out["cell_end"] = np.zeros(nc, np.int32)
        out["perm"] = np.zeros(n, np.int32)
        out["nbr_count"] = np.zeros(n, np.int32)
        if hashes:
            out["nbr_hash"] = np.zeros(n, np.uint64)
        dbg = _Debug()
        dbg.keys = out["keys"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        dbg.cell_end = out["cell_end"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        dbg.perm = out["perm"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        dbg.nbr_count = out["nbr_count"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        if hashes:
            dbg.nbr_hash = out["nbr_hash"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    rc = lib().plo_step(ctypes.byref(cs), n, pos.ctypes.data, vel.ctypes.data,
                        types.ctypes.data, ids.ctypes.data, int(nthreads),
                        ctypes.byref(dbg) if dbg is not None else None)
    if rc != 0:
        raise ValueError(f"plo_step failed with code {rc}")
    out.update(pos=pos, vel=vel, types=types, ids=ids)
    if debug:
        out["pair_count"] = int(dbg.pair_count)
        out["seconds"] = float(dbg.seconds)
    return out


def bruteforce_neighbors(settings: Settings, pos, ids=None):
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    n = pos.shape[0]
    ids = np.arange(n, dtype=np.int32) if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
    count = np.zeros(n, np.int32)
    h = np.zeros(n, np.uint64)
    cs = settings._c()
    rc = lib().plo_bruteforce_neighbors(ctypes.byref(cs), n, pos.ctypes.data, ids.ctypes.data,
                                        count.ctypes.data, h.ctypes.data)
    if rc != 0:
        raise ValueError(f"plo_bruteforce_neighbors failed with code {rc}")
    return count, h


def accelerate(a: float, beta: float, rel) -> np.ndarray:
    r = np.ascontiguousarray(rel, dtype=np.float64).reshape(3).copy()
    lib().plo_accelerate(float(a), float(beta), r.ctypes.data)
    return r


def id_hash(i: int) -> int:
    return int(lib().plo_id_hash(int(i) & 0xFFFFFFFF))


def friction_factor(friction: float, dt: float) -> float:
    return float(lib().plo_friction_factor(float(friction), float(dt)))
