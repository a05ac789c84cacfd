"""Object wrapper of a libplife batch handle (the plife_batch_* calls of include/plife.h).

A ``Batch`` steps many independent worlds in the same launches.  Each world
keeps its own settings, matrix and particle count, and its results are
bit-identical to stepping it alone in an ``Engine``.  Worlds join and leave
between steps; their ids are stable.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from ._native import PlifeSettings, PlifeTiming, PlifeTuning, check


def _settings(*, rmax, friction, dt, beta=0.3, force=1.0, wrap=True, dim=2) -> PlifeSettings:
    return PlifeSettings(int(dim), 1 if wrap else 0, float(rmax), float(friction), float(force),
                         float(dt), float(beta))


class Batch:
    def __init__(self, device: int = 0):
        self._lib = _native.load()
        h = ctypes.c_void_p()
        check(self._lib.plife_batch_create(int(device), ctypes.byref(h)))
        self._h = h
        self._n = {}                      # world id -> particle count

    # lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.plife_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # worlds
    def add(self, *, matrix, pos, vel, types, rmax, friction, dt, beta=0.3, force=1.0,
            wrap=True, dim=2) -> int:
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        s = _settings(rmax=rmax, friction=friction, dt=dt, beta=beta, force=force, wrap=wrap, dim=dim)
        wid = ctypes.c_int32()
        n = int(types.shape[0])
        check(self._lib.plife_batch_add_world(self._h, ctypes.byref(s), int(m.shape[0]), m.ctypes.data,
                                              n, pos.ctypes.data, vel.ctypes.data, types.ctypes.data,
                                              ctypes.byref(wid)))
        self._n[wid.value] = n
        return wid.value

    def add_generated(self, *, n: int, seed: int, matrix, rmax, friction, dt, beta=0.3, force=1.0,
                      wrap=True, dim=2) -> int:
        """Adds a world whose particles are generated on the device from `seed`
        (the state ``worlds.uniform_world(seed, n, ...)`` builds on the host)."""
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        s = _settings(rmax=rmax, friction=friction, dt=dt, beta=beta, force=force, wrap=wrap, dim=dim)
        wid = ctypes.c_int32()
        check(self._lib.plife_batch_add_generated_world(self._h, ctypes.byref(s), int(m.shape[0]),
                                                        m.ctypes.data, int(n),
                                                        ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                                        ctypes.byref(wid)))
        self._n[wid.value] = int(n)
        return wid.value

    def set_slots(self, wid: int, slots, pos, vel, types) -> None:
        """plife_batch_set_world_slots: overwrite particles of world wid at
        `slots` of its current order, keeping their ids."""
        slots = np.ascontiguousarray(slots, dtype=np.int32).reshape(-1)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        k = int(slots.shape[0])
        if pos.shape[0] != k or vel.shape[0] != k or types.shape[0] != k:
            raise ValueError("slots, pos, vel and types must have the same length")
        check(self._lib.plife_batch_set_world_slots(self._h, int(wid), k, slots.ctypes.data,
                                                    pos.ctypes.data, vel.ctypes.data,
                                                    types.ctypes.data))

    def set_positions(self, wid: int, preset, seed: int) -> None:
        """plife_batch_set_world_positions: a position preset for world wid."""
        from .presets import POSITION_PRESETS, preset_index
        check(self._lib.plife_batch_set_world_positions(
            self._h, int(wid), preset_index(POSITION_PRESETS, preset),
            ctypes.c_uint64(int(seed) & (2**64 - 1))))

    def set_types(self, wid: int, preset, seed: int) -> None:
        """plife_batch_set_world_types: a type preset for world wid."""
        from .presets import TYPE_PRESETS, preset_index
        check(self._lib.plife_batch_set_world_types(
            self._h, int(wid), preset_index(TYPE_PRESETS, preset),
            ctypes.c_uint64(int(seed) & (2**64 - 1))))

    def add_world(self, w) -> int:
        """Adds a ``worlds.World``."""
        return self.add(matrix=w.matrix, pos=w.pos, vel=w.vel, types=w.types, rmax=w.rmax,
                        friction=w.friction, dt=w.dt, beta=w.beta, force=w.force, wrap=w.wrap,
                        dim=w.dim)

    def remove(self, wid: int) -> None:
        check(self._lib.plife_batch_remove_world(self._h, int(wid)))
        self._n.pop(int(wid), None)

    def set_world_settings(self, wid: int, *, matrix, rmax, friction, dt, beta=0.3, force=1.0,
                           wrap=True, dim=2) -> None:
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        s = _settings(rmax=rmax, friction=friction, dt=dt, beta=beta, force=force, wrap=wrap, dim=dim)
        check(self._lib.plife_batch_set_world_settings(self._h, int(wid), ctypes.byref(s),
                                                       int(m.shape[0]), m.ctypes.data))

    def ids(self) -> list[int]:
        return sorted(self._n)

    # stepping
    def step(self, nsteps: int = 1) -> int:
        return check(self._lib.plife_batch_update(self._h, int(nsteps)))

    def stop(self) -> None:
        check(self._lib.plife_batch_force_update_stop(self._h))

    def sync(self) -> None:
        check(self._lib.plife_batch_synchronize(self._h))

    # state
    def state(self, wid: int) -> dict:
        n = self._n.get(int(wid), 0)          # an unknown id is reported by the library
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        typ = np.empty(n, np.int32)
        ids = np.empty(n, np.int32)
        check(self._lib.plife_batch_get_world(self._h, int(wid), n, pos.ctypes.data, vel.ctypes.data,
                                              typ.ctypes.data, ids.ctypes.data))
        return {"pos": pos, "vel": vel, "types": typ, "ids": ids}

    def info(self, wid: int) -> dict:
        n, pairs, steps = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.plife_batch_world_info(self._h, int(wid), ctypes.byref(n), ctypes.byref(pairs),
                                               ctypes.byref(steps)))
        return {"n": n.value, "pairs": pairs.value, "steps": steps.value}

    def count(self) -> tuple[int, int]:
        w, p = ctypes.c_int32(), ctypes.c_int64()
        check(self._lib.plife_batch_count(self._h, ctypes.byref(w), ctypes.byref(p)))
        return w.value, p.value

    # tuning and timing
    def set_tuning(self, list_cap: int = 0, stage_slack: float = 0.0,
                   direct_launch: bool = False, block_homes: int = 0,
                   sparse_worlds: int = 0, windowed_only: bool = False,
                   small_step: int = 0) -> None:
        t = PlifeTuning(int(block_homes), int(list_cap), 1 if direct_launch else 0,
                        float(stage_slack), int(sparse_worlds), 1 if windowed_only else 0,
                        int(small_step))
        check(self._lib.plife_batch_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = PlifeTuning()
        lds = ctypes.c_int32()
        check(self._lib.plife_batch_get_tuning(self._h, ctypes.byref(t), ctypes.byref(lds)))
        return {"block_homes": t.block_homes, "list_cap": t.list_cap, "stage_slack": t.stage_slack,
                "direct_launch": bool(t.direct_launch), "sparse_worlds": t.sparse_worlds,
                "windowed_only": bool(t.windowed_only), "small_step": t.small_step,
                "lds_bytes": lds.value}

    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        check(self._lib.plife_batch_time_steps(self._h, int(nsteps), ctypes.byref(ms)))
        return float(ms.value)

    def profile(self, nsteps: int) -> dict:
        t = PlifeTiming()
        check(self._lib.plife_batch_profile(self._h, int(nsteps), ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in PlifeTiming._fields_}
