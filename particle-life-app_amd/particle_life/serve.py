"""Continuous-batching server of particle-life jobs on one GPU.

A job is one world run for a given number of steps.  The reference has no
server (its app steps one world per Physics object, Physics.java:112-134);
this is the serving layer over the batch ABI: jobs are admitted in arrival
order while their device footprint fits the memory budget, join the running
batch of their dimension between updates, and leave it -- with their final
state downloaded -- as soon as they have taken their steps.  Every job's
result is bit-identical to running it alone (the batch guarantees it).

``serve_alone`` runs the same trace on the plain path -- one job at a time,
each in its own handle -- for latency comparisons.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from .batch import Batch
from .engine import Engine

# device bytes per particle: two state copies (positions u32 x3, velocities
# f32 x3, type, id) + sort scratch (keys x3, values x2, sorted keys) + world index
BYTES_PER_PARTICLE = 2 * 32 + 6 * 4 + 2
BYTES_PER_CELL = 3 * 4


def footprint(n: int, rmax: float, dim: int) -> int:
    """Device bytes a world of n particles adds to a batch."""
    nx = int(np.floor(1.0 / rmax))
    return BYTES_PER_PARTICLE * n + BYTES_PER_CELL * nx ** dim


@dataclass
class JobResult:
    job: int
    arrival: float            # s, trace clock
    admitted: float = -1.0
    done: float = -1.0
    steps: int = 0
    n: int = 0
    dim: int = 2
    state: Optional[dict] = None

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Active:
    res: JobResult
    wid: int
    left: int
    bytes: int


@dataclass
class ServeStats:
    results: list = field(default_factory=list)
    wall_s: float = 0.0
    updates: int = 0
    batch_steps: int = 0
    peak_bytes: int = 0
    host_s: dict = field(default_factory=lambda: {"admit": 0.0, "enqueue": 0.0, "wait": 0.0,
                                                   "retire": 0.0, "idle": 0.0})

    def latencies(self) -> np.ndarray:
        return np.array([r.latency for r in self.results])

    def summary(self) -> dict:
        lat = self.latencies()
        steps = sum(r.steps for r in self.results)
        psteps = sum(r.steps * r.n for r in self.results)
        return {
            "jobs": len(self.results),
            "p50_s": float(np.percentile(lat, 50)) if len(lat) else 0.0,
            "p99_s": float(np.percentile(lat, 99)) if len(lat) else 0.0,
            "mean_s": float(lat.mean()) if len(lat) else 0.0,
            "max_s": float(lat.max()) if len(lat) else 0.0,
            "wall_s": self.wall_s,
            "job_steps_per_s": steps / self.wall_s if self.wall_s else 0.0,
            "particle_steps_per_s": psteps / self.wall_s if self.wall_s else 0.0,
            "updates": self.updates,
            "peak_bytes": self.peak_bytes,
            "host_s": dict(self.host_s),
        }


def serve(jobs: list[dict], device: int = 0, budget_bytes: float = 64e9, quantum: int = 32,
          keep_state: bool = False, clock: Callable[[], float] = time.perf_counter) -> ServeStats:
    """Serves `jobs` (dicts with job, arrival (s), steps and world, a
    ``worlds.World`` holding the job's initial state) in real time.  Each
    update advances every batch by the largest step count that no running job
    overshoots (at most `quantum`); the 2D and 3D batches run on their own
    streams concurrently."""
    pending = deque(sorted(jobs, key=lambda j: j["arrival"]))
    queue: deque = deque()
    batches: dict[int, Batch] = {}
    active: dict[int, list[_Active]] = {2: [], 3: []}
    used = 0
    st = ServeStats()
    t0 = clock()
    try:
        while pending or queue or active[2] or active[3]:
            now = clock() - t0
            while pending and pending[0]["arrival"] <= now:
                queue.append(pending.popleft())
            # admission in arrival order while the footprint fits
            ta = clock()
            while queue:
                j = queue[0]
                w = j["world"]
                need = footprint(w.n, w.rmax, w.dim)
                if used + need > budget_bytes and (active[2] or active[3]):
                    break
                queue.popleft()
                b = batches.get(w.dim)
                if b is None:
                    b = batches[w.dim] = Batch(device)
                wid = b.add_world(w)
                res = JobResult(job=j["job"], arrival=j["arrival"], admitted=clock() - t0,
                                steps=j["steps"], n=w.n, dim=w.dim)
                active[w.dim].append(_Active(res, wid, j["steps"], need))
                used += need
                st.peak_bytes = max(st.peak_bytes, used)
            tb = clock()
            st.host_s["admit"] += tb - ta
            if not (active[2] or active[3]):
                if pending:                    # idle until the next arrival
                    time.sleep(max(0.0, pending[0]["arrival"] - (clock() - t0)))
                st.host_s["idle"] += clock() - tb
                continue
            stepped = []
            for dim, act in active.items():
                if not act:
                    continue
                k = min(quantum, min(a.left for a in act))
                batches[dim].step(k)
                st.batch_steps += k
                stepped.append((dim, k))
            st.updates += 1
            tc = clock()
            st.host_s["enqueue"] += tc - tb
            for dim, _ in stepped:
                batches[dim].sync()
            td = clock()
            st.host_s["wait"] += td - tc
            for dim, k in stepped:
                keep = []
                for a in active[dim]:
                    a.left -= k
                    if a.left > 0:
                        keep.append(a)
                        continue
                    if keep_state:
                        a.res.state = batches[dim].state(a.wid)
                    batches[dim].remove(a.wid)
                    a.res.done = clock() - t0
                    used -= a.bytes
                    st.results.append(a.res)
                active[dim] = keep
            st.host_s["retire"] += clock() - td
    finally:
        for b in batches.values():
            b.close()
    st.wall_s = clock() - t0
    st.results.sort(key=lambda r: r.job)
    return st


def serve_alone(jobs: list[dict], device: int = 0, keep_state: bool = False,
                clock: Callable[[], float] = time.perf_counter) -> ServeStats:
    """The plain path on the same trace: jobs run one at a time in arrival
    order, each alone in its own handle."""
    st = ServeStats()
    t0 = clock()
    e = Engine(device)
    try:
        for j in sorted(jobs, key=lambda j: j["arrival"]):
            wait = j["arrival"] - (clock() - t0)
            if wait > 0:
                time.sleep(wait)
            w = j["world"]
            res = JobResult(job=j["job"], arrival=j["arrival"], admitted=clock() - t0,
                            steps=j["steps"], n=w.n, dim=w.dim)
            e.load_world(w)
            e.step(j["steps"])
            e.sync()
            if keep_state:
                res.state = e.state()
            res.done = clock() - t0
            st.results.append(res)
            st.updates += 1
            st.batch_steps += j["steps"]
    finally:
        e.close()
    st.wall_s = clock() - t0
    st.results.sort(key=lambda r: r.job)
    return st
