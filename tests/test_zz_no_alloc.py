"""plife_step_noalloc -- the step that allocates nothing, its buffers in the
library's static device scratch -- against plife_update stepping the same
state, bit for bit, on calls in a row that shrink (the scratch keeps the last
call's contents) and change dimension.  The test's own device buffers come
from the HIP runtime the library itself uses (hipMalloc through ctypes)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime libplife.so is linked against (already loaded)."""
    from particle_life import _native

    _native.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert paths, "libamdhip64 is not loaded"
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class _Dev:
    """Device copies of host arrays, freed on close."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a: np.ndarray) -> ctypes.c_void_p:
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, n: int, dtype) -> ctypes.c_void_p:
        return self.put(np.zeros(n, dtype))

    def get(self, p, n: int, dtype) -> np.ndarray:
        out = np.zeros(n, dtype)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _noalloc_step(hip, w):
    from particle_life import _native

    lib = _native.load()
    n, d3 = w.n, w.dim == 3
    fx = (np.round(w.pos * 2.0 ** 32).astype(np.uint64) % (1 << 32)).astype(np.uint32)
    dev = _Dev(hip)
    try:
        inp = {}
        for k, ax in enumerate("xyz"):
            inp["p" + ax] = dev.put(fx[:, k].copy())
            inp["v" + ax] = dev.put(w.vel[:, k].astype(np.float32))
        inp["type"] = dev.put(np.asarray(w.types, np.int32))
        inp["id"] = dev.put(np.arange(n, dtype=np.int32))
        out = {k: dev.empty(n, np.uint32 if k[0] == "p" else (np.float32 if k[0] == "v" else np.int32))
               for k in inp}
        mat = dev.put(np.ascontiguousarray(w.matrix, dtype=np.float64))
        pairs = dev.empty(1, np.uint64)
        s = _native.PlifeSettings(int(w.dim), 1 if w.wrap else 0, float(w.rmax), float(w.friction),
                                  float(w.force), float(w.dt), float(w.beta))

        def z(key):          # z arrays only in 3D
            return inp[key] if d3 else None

        def zo(key):
            return out[key] if d3 else None

        rc = lib.plife_step_noalloc(
            ctypes.byref(s), int(w.matrix.shape[0]), mat, n,
            inp["px"], inp["py"], z("pz"), inp["vx"], inp["vy"], z("vz"), inp["type"], inp["id"],
            out["px"], out["py"], zo("pz"), out["vx"], out["vy"], zo("vz"), out["type"], out["id"],
            pairs, None)
        _native.check(rc)
        assert hip.hipDeviceSynchronize() == 0
        pos = np.zeros((n, 3))
        vel = np.zeros((n, 3))
        for k, ax in enumerate("xyz"[:w.dim]):
            pos[:, k] = dev.get(out["p" + ax], n, np.uint32).astype(np.float64) * 2.0 ** -32
            vel[:, k] = dev.get(out["v" + ax], n, np.float32).astype(np.float64)
        return {"pos": pos, "vel": vel, "types": dev.get(out["type"], n, np.int32),
                "ids": dev.get(out["id"], n, np.int32), "pairs": int(dev.get(pairs, 1, np.uint64)[0])}
    finally:
        dev.close()


def test_noalloc_step_matches_update_on_shrinking_calls():
    from particle_life import Engine, worlds

    hip = _hip()
    worlds_in_turn = [
        worlds.uniform_world(7001, 20_000, 6, rmax=0.02, dt=0.02),
        worlds.uniform_world(7002, 5_000, 4, rmax=0.035, dt=0.01, wrap=False),
        worlds.uniform_world(7003, 3_000, 5, rmax=0.09, dt=0.01, dim=3),
    ]
    for k, w in enumerate(worlds_in_turn):
        print(f"\n[noalloc] call {k}: n={w.n} dim={w.dim} wrap={w.wrap} rmax={w.rmax}", flush=True)
        got = _noalloc_step(hip, w)
        with Engine(0) as e:
            e.load_world(w)
            e.step(1)
            ref = e.state()
            ref_pairs = e.pair_count()
        for key in ("pos", "vel", "types", "ids"):
            np.testing.assert_array_equal(got[key], ref[key], err_msg=f"call {k}: {key}")
        assert got["pairs"] == ref_pairs, (got["pairs"], ref_pairs)
        assert ref_pairs > 0


def test_noalloc_step_rejects_sizes_beyond_its_scratch():
    from particle_life import _native

    lib = _native.load()
    s = _native.PlifeSettings(2, 1, 0.02, 0.85, 1.0, 0.02, 0.3)
    m = (ctypes.c_double * 1)(0.5)
    rc = lib.plife_step_noalloc(ctypes.byref(s), 1, ctypes.cast(m, ctypes.c_void_p), 262_145,
                                *([None] * 16), None, None)
    assert rc == _native.PLIFE_ERR_ARGUMENT
    assert b"scratch limit" in lib.plife_last_error()
    s3 = _native.PlifeSettings(3, 1, 0.005, 0.85, 1.0, 0.02, 0.3)      # 200^3 cells
    rc = lib.plife_step_noalloc(ctypes.byref(s3), 1, ctypes.cast(m, ctypes.c_void_p), 10,
                                *([None] * 16), None, None)
    assert rc == _native.PLIFE_ERR_ARGUMENT
