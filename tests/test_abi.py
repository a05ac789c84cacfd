"""The C-ABI library loads and exports exactly what include/plife.h declares
(no compute: these run without a GPU)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from particle_life import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "plife.h")).read()
    return re.findall(r"^\s*(?:int|const char \*)\s*(plife_\w+)\s*\(", src, re.M)


def test_every_declared_symbol_is_exported():
    lib = _native.load()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(names) == sorted(_native.EXPORTS)


def test_library_is_gfx950_hip_build():
    path = _native.LIB_PATH
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


def test_null_arguments_are_rejected_without_a_device():
    lib = _native.load()
    s = _native.PlifeSettings(2, 1, 0.02, 0.85, 1.0, 0.02, 0.3)
    assert lib.plife_set_settings(None, ctypes.byref(s), 1, None) == _native.PLIFE_ERR_ARGUMENT
    assert b"null" in lib.plife_last_error()
    assert lib.plife_update(None, 1) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_create(0, None) == _native.PLIFE_ERR_ARGUMENT
    with pytest.raises(_native.PlifeArgumentError):
        _native.check(lib.plife_update(None, 1))


def test_status_codes_map_to_java_exceptions():
    assert issubclass(_native.PlifeArgumentError, ValueError)
    assert issubclass(_native.PlifeError, RuntimeError)


def test_batch_calls_reject_null_handles_without_a_device():
    lib = _native.load()
    s = _native.PlifeSettings(2, 1, 0.02, 0.85, 1.0, 0.02, 0.3)
    wid = ctypes.c_int32()
    assert lib.plife_batch_create(0, None) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_batch_add_world(None, ctypes.byref(s), 1, None, 0, None, None, None,
                                     ctypes.byref(wid)) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_batch_update(None, 1) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_batch_remove_world(None, 0) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_batch_get_world(None, 0, 0, None, None, None, None) == _native.PLIFE_ERR_ARGUMENT
    assert lib.plife_batch_destroy(None) == _native.PLIFE_OK
