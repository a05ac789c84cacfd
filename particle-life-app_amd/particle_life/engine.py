"""Thin object wrapper of one libplife handle (include/plife.h).

``Engine`` is the direct, array-in/array-out form of the C ABI used by the
tests and the benchmark; ``GpuPhysics`` (physics.py) layers the reference's
Physics API on top of it.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _native
from ._native import PlifeSettings, PlifeTiming, PlifeTuning, check


class Engine:
    def __init__(self, device: int = 0):
        self._lib = _native.load()
        h = ctypes.c_void_p()
        check(self._lib.plife_create(int(device), ctypes.byref(h)))
        self._h = h
        self.n = 0

    # lifecycle
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.plife_destroy(self._h)
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

    # configuration and state
    def configure(self, *, matrix, rmax: float, friction: float, dt: float, beta: float = 0.3,
                  force: float = 1.0, wrap: bool = True, dim: int = 2) -> None:
        m = np.ascontiguousarray(matrix, dtype=np.float64)
        s = PlifeSettings(int(dim), 1 if wrap else 0, float(rmax), float(friction), float(force),
                          float(dt), float(beta))
        check(self._lib.plife_set_settings(self._h, ctypes.byref(s), int(m.shape[0]), m.ctypes.data))

    def configure_world(self, w) -> None:
        self.configure(matrix=w.matrix, rmax=w.rmax, friction=w.friction, dt=w.dt, beta=w.beta,
                       force=w.force, wrap=w.wrap, dim=w.dim)

    def set_state(self, pos, vel, types) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        n = int(types.shape[0])
        check(self._lib.plife_set_particles(self._h, n, pos.ctypes.data, vel.ctypes.data,
                                            types.ctypes.data))
        self.n = n

    def generate(self, n: int, seed: int) -> None:
        """Replaces the particles with n generated on the device from `seed`
        (uniform positions and types, v = 0): the same state as
        ``worlds.uniform_world(seed, n, ...)`` builds on the host."""
        check(self._lib.plife_generate_particles(self._h, int(n), ctypes.c_uint64(int(seed) & (2**64 - 1))))
        self.n = int(n)

    def set_slots(self, slots, pos, vel, types) -> None:
        """plife_set_particle_slots: overwrite the particles at `slots` of the
        current order (the order state() returns), keeping their ids."""
        slots = np.ascontiguousarray(slots, dtype=np.int32).reshape(-1)
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        types = np.ascontiguousarray(types, dtype=np.int32).reshape(-1)
        k = int(slots.shape[0])
        if pos.shape[0] != k or vel.shape[0] != k or types.shape[0] != k:
            raise ValueError("slots, pos, vel and types must have the same length")
        check(self._lib.plife_set_particle_slots(self._h, k, slots.ctypes.data, pos.ctypes.data,
                                                 vel.ctypes.data, types.ctypes.data))

    def set_positions(self, preset, seed: int) -> None:
        """plife_set_positions: a position preset (name or index, the app's
        order) for every particle, velocities 0, types kept."""
        from .presets import POSITION_PRESETS, preset_index
        check(self._lib.plife_set_positions(self._h, preset_index(POSITION_PRESETS, preset),
                                            ctypes.c_uint64(int(seed) & (2**64 - 1))))

    def set_types(self, preset, seed: int) -> None:
        """plife_set_types: a type preset (name or index, the app's order)."""
        from .presets import TYPE_PRESETS, preset_index
        check(self._lib.plife_set_types(self._h, preset_index(TYPE_PRESETS, preset),
                                        ctypes.c_uint64(int(seed) & (2**64 - 1))))

    def load_world(self, w) -> None:
        self.configure_world(w)
        self.set_state(w.pos, w.vel, w.types)

    def state(self) -> dict:
        n = self.n
        pos = np.empty((n, 3))
        vel = np.empty((n, 3))
        typ = np.empty(n, np.int32)
        ids = np.empty(n, np.int32)
        self.download(pos, vel, typ, ids)
        return {"pos": pos, "vel": vel, "types": typ, "ids": ids}

    def download(self, pos=None, vel=None, types=None, ids=None) -> None:
        """plife_get_particles into caller-owned arrays (None = not wanted):
        pos/vel C-contiguous float64 (n, 3), types/ids C-contiguous int32 (n,).
        Reusing the arrays across frames saves the page faults of fresh ones."""
        n = self.n
        ptrs = []
        for a, shape, dt in ((pos, (n, 3), np.float64), (vel, (n, 3), np.float64),
                             (types, (n,), np.int32), (ids, (n,), np.int32)):
            if a is None:
                ptrs.append(None)
                continue
            if a.shape != shape or a.dtype != dt or not a.flags.c_contiguous or not a.flags.writeable:
                raise ValueError(f"download needs a writeable C-contiguous {np.dtype(dt).name} "
                                 f"array of shape {shape}, got {a.dtype} {a.shape}")
            ptrs.append(a.ctypes.data)
        check(self._lib.plife_get_particles(self._h, n, *ptrs))

    # stepping
    def step(self, nsteps: int = 1) -> int:
        return check(self._lib.plife_update(self._h, int(nsteps)))

    def stop(self) -> None:
        check(self._lib.plife_force_update_stop(self._h))

    def steps_taken(self) -> int:
        """Steps queued on this handle since it was created (each completes)."""
        v = ctypes.c_int64()
        check(self._lib.plife_get_steps(self._h, ctypes.byref(v)))
        return int(v.value)

    def sync(self) -> None:
        check(self._lib.plife_synchronize(self._h))

    # parity exports
    def grid(self) -> tuple[int, int]:
        nx, nc = ctypes.c_int32(), ctypes.c_int32()
        check(self._lib.plife_get_grid(self._h, ctypes.byref(nx), ctypes.byref(nc)))
        return nx.value, nc.value

    def cells(self) -> dict:
        _, nc = self.grid()
        keys = np.zeros(self.n, np.int32)
        cell_end = np.zeros(nc, np.int32)
        perm = np.zeros(self.n, np.int32)
        check(self._lib.plife_get_cells(self._h, keys.ctypes.data, cell_end.ctypes.data,
                                        perm.ctypes.data, nc))
        return {"keys": keys, "cell_end": cell_end, "perm": perm}

    def pair_count(self) -> int:
        v = ctypes.c_int64()
        check(self._lib.plife_get_pair_count(self._h, ctypes.byref(v)))
        return int(v.value)

    def set_debug(self, on: bool) -> None:
        check(self._lib.plife_set_debug(self._h, 1 if on else 0))

    def neighbor_summary(self) -> tuple[np.ndarray, np.ndarray]:
        count = np.zeros(self.n, np.int32)
        h = np.zeros(self.n, np.uint64)
        check(self._lib.plife_get_neighbor_summary(self._h, self.n, count.ctypes.data, h.ctypes.data))
        return count, h

    def accelerations(self) -> np.ndarray:
        """Debug mode: each particle's acceleration of the last step, FP32 (n, 3)."""
        acc = np.zeros((self.n, 3), np.float32)
        check(self._lib.plife_get_accelerations(self._h, self.n, acc.ctypes.data))
        return acc

    # launch tuning (results do not depend on it)
    def set_tuning(self, list_cap: int = 0, stage_slack: float = 0.0,
                   direct_launch: bool = False, block_homes: int = 0,
                   sparse_worlds: int = 0, windowed_only: bool = False,
                   small_step: int = 0) -> None:
        t = PlifeTuning(int(block_homes), int(list_cap), 1 if direct_launch else 0,
                        float(stage_slack), int(sparse_worlds), 1 if windowed_only else 0,
                        int(small_step))
        check(self._lib.plife_set_tuning(self._h, ctypes.byref(t)))

    def tuning(self) -> dict:
        t = PlifeTuning()
        lds = ctypes.c_int32()
        check(self._lib.plife_get_tuning(self._h, ctypes.byref(t), ctypes.byref(lds)))
        return {"block_homes": t.block_homes, "list_cap": t.list_cap, "stage_slack": t.stage_slack,
                "direct_launch": bool(t.direct_launch), "sparse_worlds": t.sparse_worlds,
                "windowed_only": bool(t.windowed_only), "small_step": t.small_step,
                "lds_bytes": lds.value}

    # timing
    def time_steps(self, nsteps: int) -> float:
        ms = ctypes.c_double()
        check(self._lib.plife_time_steps(self._h, int(nsteps), ctypes.byref(ms)))
        return float(ms.value)

    def profile(self, nsteps: int) -> dict:
        t = PlifeTiming()
        check(self._lib.plife_profile(self._h, int(nsteps), ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in PlifeTiming._fields_}
