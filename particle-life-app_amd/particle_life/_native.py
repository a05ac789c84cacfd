"""ctypes binding of libplife.so (the C ABI declared in include/plife.h).

The library is built in-tree by ``__graft_entry__.build()``.  There is no
fallback: if the shared object is missing or fails to load, importing the
binding raises, so nothing can silently run a CPU path instead of the HIP one.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplife.so")

PLIFE_OK = 0
PLIFE_STOPPED = 1
PLIFE_ERR_ARGUMENT = -1
PLIFE_ERR_STATE = -2
PLIFE_ERR_DEVICE = -3
PLIFE_ERR_MEMORY = -4

# Every symbol include/plife.h declares, in declaration order.
EXPORTS = (
    "plife_create",
    "plife_destroy",
    "plife_set_settings",
    "plife_set_particles",
    "plife_set_particle_slots",
    "plife_get_particles",
    "plife_get_count",
    "plife_generate_particles",
    "plife_set_positions",
    "plife_set_types",
    "plife_generate_matrix",
    "plife_update",
    "plife_force_update_stop",
    "plife_get_steps",
    "plife_synchronize",
    "plife_last_error",
    "plife_get_cells",
    "plife_get_grid",
    "plife_get_pair_count",
    "plife_set_debug",
    "plife_get_neighbor_summary",
    "plife_get_accelerations",
    "plife_step_noalloc",
    "plife_set_tuning",
    "plife_get_tuning",
    "plife_profile",
    "plife_time_steps",
    "plife_batch_create",
    "plife_batch_destroy",
    "plife_batch_add_world",
    "plife_batch_set_world_slots",
    "plife_batch_set_world_positions",
    "plife_batch_set_world_types",
    "plife_batch_add_generated_world",
    "plife_batch_remove_world",
    "plife_batch_set_world_settings",
    "plife_batch_update",
    "plife_batch_force_update_stop",
    "plife_batch_synchronize",
    "plife_batch_get_world",
    "plife_batch_world_info",
    "plife_batch_count",
    "plife_batch_set_tuning",
    "plife_batch_get_tuning",
    "plife_batch_time_steps",
    "plife_batch_profile",
)


class PlifeSettings(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32),
        ("wrap", ctypes.c_int32),
        ("rmax", ctypes.c_double),
        ("friction", ctypes.c_double),
        ("force", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("beta", ctypes.c_double),
    ]


class PlifeTiming(ctypes.Structure):
    _fields_ = [
        ("step_ms", ctypes.c_double),
        ("bin_ms", ctypes.c_double),
        ("sort_ms", ctypes.c_double),
        ("gather_ms", ctypes.c_double),
        ("force_ms", ctypes.c_double),
        ("pairs", ctypes.c_int64),
        ("steps", ctypes.c_int32),
    ]


class PlifeTuning(ctypes.Structure):
    _fields_ = [
        ("block_homes", ctypes.c_int32),
        ("list_cap", ctypes.c_int32),
        ("direct_launch", ctypes.c_int32),
        ("stage_slack", ctypes.c_double),
        ("sparse_worlds", ctypes.c_int32),
        ("windowed_only", ctypes.c_int32),
        ("small_step", ctypes.c_int32),
    ]


class PlifeError(RuntimeError):
    """A failure reported by the library (Java side: RuntimeException)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"plife error {code}: {message}")
        self.code = code


class PlifeArgumentError(PlifeError, ValueError):
    """Bad argument (Java side: IllegalArgumentException)."""


_lib = None


def load() -> ctypes.CDLL:
    """Load libplife.so, raising if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64p = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64)
    sig = {
        "plife_create": [i32, ctypes.POINTER(vp)],
        "plife_destroy": [vp],
        "plife_set_settings": [vp, ctypes.POINTER(PlifeSettings), i32, vp],
        "plife_set_particles": [vp, i32, vp, vp, vp],
        "plife_set_particle_slots": [vp, i32, vp, vp, vp, vp],
        "plife_get_particles": [vp, i32, vp, vp, vp, vp],
        "plife_get_count": [vp, ctypes.POINTER(i32)],
        "plife_generate_particles": [vp, i32, ctypes.c_uint64],
        "plife_set_positions": [vp, i32, ctypes.c_uint64],
        "plife_set_types": [vp, i32, ctypes.c_uint64],
        "plife_generate_matrix": [i32, i32, ctypes.c_uint64, vp],
        "plife_update": [vp, i32],
        "plife_force_update_stop": [vp],
        "plife_get_steps": [vp, i64p],
        "plife_synchronize": [vp],
        "plife_get_cells": [vp, vp, vp, vp, i32],
        "plife_get_grid": [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)],
        "plife_get_pair_count": [vp, i64p],
        "plife_set_debug": [vp, i32],
        "plife_get_neighbor_summary": [vp, i32, vp, vp],
        "plife_get_accelerations": [vp, i32, vp],
        "plife_step_noalloc": [ctypes.POINTER(PlifeSettings), i32, vp, i32] + [vp] * 16 + [vp, vp],
        "plife_set_tuning": [vp, ctypes.POINTER(PlifeTuning)],
        "plife_get_tuning": [vp, ctypes.POINTER(PlifeTuning), ctypes.POINTER(i32)],
        "plife_profile": [vp, i32, ctypes.POINTER(PlifeTiming)],
        "plife_time_steps": [vp, i32, ctypes.POINTER(ctypes.c_double)],
        "plife_batch_create": [i32, ctypes.POINTER(vp)],
        "plife_batch_destroy": [vp],
        "plife_batch_add_world": [vp, ctypes.POINTER(PlifeSettings), i32, vp, i32, vp, vp, vp,
                                  ctypes.POINTER(i32)],
        "plife_batch_set_world_slots": [vp, i32, i32, vp, vp, vp, vp],
        "plife_batch_set_world_positions": [vp, i32, i32, ctypes.c_uint64],
        "plife_batch_set_world_types": [vp, i32, i32, ctypes.c_uint64],
        "plife_batch_add_generated_world": [vp, ctypes.POINTER(PlifeSettings), i32, vp, i32,
                                            ctypes.c_uint64, ctypes.POINTER(i32)],
        "plife_batch_remove_world": [vp, i32],
        "plife_batch_set_world_settings": [vp, i32, ctypes.POINTER(PlifeSettings), i32, vp],
        "plife_batch_update": [vp, i32],
        "plife_batch_force_update_stop": [vp],
        "plife_batch_synchronize": [vp],
        "plife_batch_get_world": [vp, i32, i32, vp, vp, vp, vp],
        "plife_batch_world_info": [vp, i32, ctypes.POINTER(i32), i64p, i64p],
        "plife_batch_count": [vp, ctypes.POINTER(i32), i64p],
        "plife_batch_set_tuning": [vp, ctypes.POINTER(PlifeTuning)],
        "plife_batch_get_tuning": [vp, ctypes.POINTER(PlifeTuning), ctypes.POINTER(i32)],
        "plife_batch_time_steps": [vp, i32, ctypes.POINTER(ctypes.c_double)],
        "plife_batch_profile": [vp, i32, ctypes.POINTER(PlifeTiming)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.plife_last_error.argtypes = []
    L.plife_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int) -> int:
    """Map a status code to the exception the Java layer would throw."""
    if rc >= 0:
        return rc
    msg = (load().plife_last_error() or b"").decode("utf-8", "replace")
    if rc == PLIFE_ERR_ARGUMENT:
        raise PlifeArgumentError(rc, msg)
    raise PlifeError(rc, msg)
