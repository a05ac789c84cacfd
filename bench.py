#!/usr/bin/env python3
"""Benchmark of the particle-life physics step (Physics.update()) on one MI355X.

Workload (default, --config c2): the single large 2D world -- N = 1,000,000
particles, 8 types, rmax 0.01 in [-1,1)^2 (0.005 in the reference's unit box,
~78.5 interacting neighbours), beta 0.3, friction half-life 0.04 s, dt 0.01 s,
periodic.  One step = one Physics.update(): cell keys, stable sort, fused
force + integrate.  --config c1 / c3 run the 10k 2D and clustered 4M 3D worlds.

Prints ONE JSON line.  `value` is pair-interactions/s (ordered interacting
pairs per step, counted exactly by the kernel, times steps/s) with the state
resident in HBM; the timed region is K steps between two HIP events on the
library's stream, bracketed by synchronisations.

roofline: the force+integrate kernel (the dominant one) against the FP32
vector peak, with algorithmic work of 22 FLOP per interacting 2D pair (28 in 3D).
cpu_baseline: the FP64 C restatement (oracle/) on this host's cores, a bounded
sample of the same world.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "particle-life-app_amd"))
sys.path.insert(0, ROOT)

FLOP_PER_PAIR = {2: 22, 3: 28}          # algorithmic FLOPs per interacting pair
BYTES_PER_PARTICLE = {2: 36, 3: 52}      # read x,v,type + write x,v
PEAK_FP32_TFLOPS = 157.3       # MI355X FP32 vector peak
PEAK_HBM_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=24)
    p.add_argument("--config", choices=("c1", "c2", "c3", "c4", "c5"), default="c2",
                   help="c2 (default): the 1M-particle 2D world the metric is quoted on; "
                        "c1: the 10k 2D world; c3: the clustered 4M 3D world; "
                        "c4: 512 independent 2D worlds stepped as one batch; "
                        "c5: a Poisson stream of jobs served by continuous batching")
    p.add_argument("--jobs", type=int, default=200, help="c5: jobs in the trace")
    p.add_argument("--rate", type=float, default=40.0, help="c5: arrivals per second")
    p.add_argument("--quantum", type=int, default=32, help="c5: most steps per batch update")
    p.add_argument("--worlds", type=int, default=512, help="c4: number of worlds")
    p.add_argument("--no-plain", action="store_true",
                   help="c4/c5: skip the comparison with the plain (one world per handle) path")
    p.add_argument("--n", type=int, default=0, help="particle count override (0 = the config's)")
    p.add_argument("--profile-steps", type=int, default=20)
    p.add_argument("--cpu-steps", type=int, default=0,
                   help="steps of the CPU sample; 0 = 100 (c1), 10 (c2), 3 (c3)")
    p.add_argument("--cpu-threads", type=int, default=0,
                   help="0 = every available CPU (affinity and cgroup quota), at most 16")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--tuning", default="",
                   help="c1-c3: launch tuning as JSON, e.g. '{\"windowed_only\": true}' "
                        "(fields of plife_tuning; results do not depend on it)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                   help="instead of the benchmark, run 5 plain device copies of BYTES bytes "
                        "(a known-size kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def copy_calibration(nbytes: int) -> int:
    import torch

    a = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}))
    return 0


def force_kernel_sha256() -> str:
    """Hash of the force kernel's sources (the build the counters belong to)."""
    import hashlib

    h = hashlib.sha256()
    for name in ("plife_force.hpp", "plife_kernels.hpp"):
        with open(os.path.join(ROOT, "particle-life-app_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def host_cpus() -> dict:
    """What the host offers the CPU baseline: logical CPUs, those this process
    may run on (affinity), the cgroup's CPU quota, and the CPU model."""
    info = {"nproc": os.cpu_count() or 1}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = info["nproc"]
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    info["cgroup_cpus"] = quota
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    info["model"] = model
    avail = info["affinity"]
    if quota:
        avail = min(avail, max(1, int(quota + 0.5)))
    info["available"] = avail
    return info


# a GPU box grants one GPU's process this many CPU threads at most
CPU_SHARE = 16


def cpu_threads(requested: int, host: dict) -> int:
    if requested > 0:
        return requested
    return max(1, min(CPU_SHARE, host["available"]))


def cpu_baseline(world, steps: int, threads: int, host: dict) -> dict:
    """FP64 restatement of Physics.update() on the host, same world.  The time
    is the step proper as plo_step measures it inside (binning and the two
    parallel phases over `threads` threads), without the wrapper's copies in
    and out of its particle records."""
    import oracle

    S = oracle.Settings(matrix=world.matrix, rmax=world.rmax, friction=world.friction,
                        dt=world.dt, beta=world.beta, force=world.force, wrap=world.wrap,
                        dim=world.dim)
    pos, vel, typ = world.pos, world.vel, world.types
    ids = None
    pairs = 0
    t_total = 0.0
    for _ in range(steps):
        r = oracle.step(S, pos, vel, typ, ids=ids, nthreads=threads, debug=True, hashes=False)
        t_total += r["seconds"]
        pairs += r["pair_count"]
        pos, vel, typ, ids = r["pos"], r["vel"], r["types"], r["ids"]
    return {
        "value": pairs / t_total,
        "unit": "pair-interactions/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the same N={world.n} world from its initial state, "
                  f"FP64 C restatement, {threads} threads, timed inside plo_step "
                  f"({t_total:.2f} s)",
        "steps_per_s": steps / t_total,
        "host": host,
    }


def bench_batch(args) -> int:
    """512 independent worlds (N log-uniform in [1000, 65536]) stepped as one
    batch, against the plain path: one handle per world, each on its own
    stream, stepped round-robin from the host."""
    from particle_life import Batch, Engine, worlds

    ws = worlds.config4(args.worlds)
    ntot = sum(w.n for w in ws)
    b = Batch(0)
    ids = [b.add_world(w) for w in ws]
    b.step(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    ms = b.time_steps(args.steps)
    wall = time.perf_counter() - t0
    pairs = sum(b.info(i)["pairs"] for i in ids)
    launch = b.tuning()
    prof = b.profile(args.profile_steps)
    b.close()
    steps_per_s = args.steps / (ms / 1e3)
    out = {
        "metric": f"pair-interactions/s & world-steps/s ({args.worlds}-world batch)",
        "value": pairs * steps_per_s,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "steps_per_s": steps_per_s,
        "world_steps_per_s": args.worlds * steps_per_s,
        "particle_steps_per_s": ntot * steps_per_s,
        "pairs_per_step": pairs,
        "wall_s": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 force law, u32 fixed-point positions",
        "data": "synthetic (seeded SplitMix64 worlds)",
        "config": {
            "workload": f"{args.worlds} independent 2D worlds, N log-uniform in [1000, 65536] "
                        f"(total {ntot}), 3-12 types, rmax U[0.01, 0.04] (unit box), "
                        "beta U[0.2, 0.4], half-life 0.04 s, dt 0.01 s, periodic",
            "n": ntot, "worlds": args.worlds,
        },
        "launch": {k: launch[k] for k in ("block_homes", "list_cap", "small_step", "lds_bytes")},
        "phases_ms": {k: prof[k] for k in ("bin_ms", "sort_ms", "gather_ms", "force_ms", "step_ms")},
    }
    if not args.no_plain:
        # the plain path: one handle (and stream) per world
        engines = []
        for w in ws:
            e = Engine(0)
            e.load_world(w)
            engines.append(e)
        k2 = max(2, args.steps // 2 * 2)
        for e in engines:
            e.step(2)
        for e in engines:
            e.sync()
        t0 = time.perf_counter()
        for _ in range(k2 // 2):
            for e in engines:
                e.step(2)
        for e in engines:
            e.sync()
        plain_s = time.perf_counter() - t0
        for e in engines:
            e.close()
        out["plain_path"] = {
            "ms_per_step": plain_s * 1e3 / k2,
            "world_steps_per_s": args.worlds * k2 / plain_s,
            "steps": k2,
            "how": "one handle per world (own stream, HIP graph of 2 steps), stepped round-robin",
        }
        out["speedup_vs_plain"] = out["world_steps_per_s"] / out["plain_path"]["world_steps_per_s"]
    print(json.dumps(out))
    return 0


def bench_serve(args) -> int:
    """A Poisson stream of jobs (2D and 3D worlds of 2k-262k particles run for
    100-2000 steps) served by continuous batching, against the plain path
    serving the same trace one job at a time."""
    from particle_life import worlds
    from particle_life.serve import serve, serve_alone

    jobs = worlds.config5_jobs(args.jobs, rate=args.rate)
    for j in jobs:
        j["world"] = worlds.job_world(j)          # clients bring their initial state
    served = serve(jobs, quantum=args.quantum)
    s = served.summary()
    out = {
        "metric": "served job latency p99 (continuous batching)",
        "value": s["p99_s"],
        "unit": "s",
        "n_gpus": 1,
        "p50_s": s["p50_s"],
        "p99_s": s["p99_s"],
        "mean_s": s["mean_s"],
        "jobs": s["jobs"],
        "wall_s": s["wall_s"],
        "job_steps_per_s": s["job_steps_per_s"],
        "particle_steps_per_s": s["particle_steps_per_s"],
        "updates": s["updates"],
        "host_s": s["host_s"],
        "quantum": args.quantum,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 force law, u32 fixed-point positions",
        "data": "synthetic (seeded SplitMix64 job trace)",
        "config": {
            "workload": f"{args.jobs} jobs, Poisson arrivals at {args.rate}/s; each a 2D or 3D "
                        "periodic world, N log-uniform in [2000, 262144], 4-10 types, rmax "
                        "U[0.01, 0.03] (unit box), beta 0.3, dt 0.01 s, 100-2000 steps; "
                        "64 GB admission budget",
            "jobs": args.jobs, "rate": args.rate,
        },
    }
    if not args.no_plain:
        alone = serve_alone(jobs).summary()
        out["plain_path"] = {k: alone[k] for k in ("p50_s", "p99_s", "mean_s", "wall_s",
                                                   "job_steps_per_s", "particle_steps_per_s")}
        out["plain_path"]["how"] = "jobs one at a time in arrival order, each alone in one handle"
        out["p99_speedup_vs_plain"] = alone["p99_s"] / s["p99_s"]
        out["p50_speedup_vs_plain"] = alone["p50_s"] / s["p50_s"]
    print(json.dumps(out))
    return 0


def main() -> int:
    args = parse()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the world is stepped "
              "on a single MI355X and multi-GPU sharding is not part of this build.",
              file=sys.stderr)
        return 2
    if args.config == "c4":
        return bench_batch(args)
    if args.config == "c5":
        return bench_serve(args)

    import numpy as np

    from particle_life import Engine, worlds

    make = {"c1": worlds.config1, "c2": worlds.config2, "c3": worlds.config3}[args.config]
    w = make(args.n) if args.n else make()
    eng = Engine(0)
    if args.tuning:
        eng.set_tuning(**json.loads(args.tuning))
    eng.load_world(w)
    eng.step(args.warmup)
    eng.sync()

    # timed region: K steps between two events on the library's stream
    eng.sync()
    t0 = time.perf_counter()
    ms = eng.time_steps(args.steps)
    eng.sync()
    wall = time.perf_counter() - t0
    pairs_last = eng.pair_count()
    launch = eng.tuning()
    one_launch = launch["small_step"] == 2

    # per-phase timing (events around each phase) on further steps
    prof = eng.profile(args.profile_steps)

    # the app's path: one update and the snapshot's download per frame
    # (PhysicsSnapshot.write: positions, velocities, types into arrays kept
    # across frames), host-visible wall time
    frames = 20
    snap_pos, snap_vel = np.empty((w.n, 3)), np.empty((w.n, 3))
    snap_typ = np.empty(w.n, np.int32)
    for _ in range(2):          # the first download allocates the staging buffers
        eng.step(1)
        eng.download(snap_pos, snap_vel, snap_typ)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(frames):
        eng.step(1)
        eng.download(snap_pos, snap_vel, snap_typ)
    frame_ms = (time.perf_counter() - t0) * 1e3 / frames
    t0 = time.perf_counter()
    for _ in range(frames):
        eng.download(snap_pos, snap_vel, snap_typ)
    get_state_ms = (time.perf_counter() - t0) * 1e3 / frames
    pairs = prof["pairs"]
    steps_per_s = args.steps / (ms / 1e3)
    value = pairs_last * steps_per_s

    force_s = prof["force_ms"] / 1e3
    achieved_tflops = FLOP_PER_PAIR[w.dim] * pairs / force_s / 1e12

    # HBM bytes per force launch from rocprofv3 FETCH_SIZE / WRITE_SIZE, used
    # only when they were measured on this build of the force kernel
    traffic, traffic_note = None, "not measured for this workload"
    tpath = os.path.join(ROOT, "profiles", "force_c2_traffic.json")
    if args.config == "c2" and not args.n and os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f)
        if rec.get("kernel_source_sha256") == force_kernel_sha256():
            traffic = rec.get("bytes_per_launch")
            traffic_note = "rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE x 1 (scales from a 256 MiB copy), " \
                           "profiles/force_c2_counters.md"
        else:
            traffic_note = "the recorded counters belong to another build of the force kernel"


    workloads = {
        "c1": "plain 2D world: N=10,000, 6 types, rmax 0.04 in [-1,1)^2 (0.02 unit box), beta 0.3, "
              "friction half-life 0.04 s, dt 0.02 s, wrap",
        "c2": "single large 2D world: N=1,000,000, 8 types, rmax 0.01 in [-1,1)^2 "
              "(0.005 unit box), beta 0.3, friction half-life 0.04 s, dt 0.01 s, wrap",
        "c3": "clustered 3D world: N=4,000,000, 12 types, 64 Gaussian blobs (sigma 0.03 in "
              "[-1,1)^3), rmax 0.02 (0.01 unit box), beta 0.25, dt 0.01 s, wrap",
    }
    out = {
        "metric": "pair-interactions/s & steps/s at N=1M 2D" if args.config == "c2"
                  else f"pair-interactions/s & steps/s ({args.config})",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "steps_per_s": steps_per_s,
        "particle_steps_per_s": w.n * steps_per_s,
        "pairs_per_step": pairs_last,
        "wall_s": wall,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 force law, u32 fixed-point positions",
        "data": "synthetic (seeded SplitMix64 world)",
        "config": {
            "workload": workloads[args.config],
            "n": w.n, "dim": w.dim, "ntypes": w.ntypes, "rmax_unit_box": w.rmax, "dt": w.dt,
        },
        "step_path": ("one launch per step (k_step_small: keys, sort, cell table and force in one "
                      "workgroup's LDS)" if one_launch else
                      "multi-launch (bin, scans, radix sort, force)"),
        "launch": {k: launch[k] for k in ("block_homes", "list_cap", "small_step", "lds_bytes")},
        "phases_ms": {k: prof[k] for k in ("bin_ms", "sort_ms", "gather_ms", "force_ms", "step_ms")},
        "app_path": {"ms_per_frame": frame_ms, "get_state_ms": get_state_ms,
                     "how": "update(1) + plife_get_particles of positions, velocities and types "
                            "into double[3N]/int[N] arrays reused across frames (as "
                            "PhysicsSnapshot.write), wall clock"},
        "roofline": {
            "kernel": (f"k_step_small<{w.dim}, no-debug> (the whole step)" if one_launch
                       else f"k_force<{w.dim}, wrap, no-debug>"),
            "bound": "valu",            # vector ALU issue: sparse gather, no matrix ops
            "pipe": "FP32 screen + FP64 force law as vector instructions; priced against the "
                    "FP32 vector peak",
            "achieved": achieved_tflops,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_note": traffic_note,
            "algorithmic": f"{FLOP_PER_PAIR[w.dim]} FLOP/pair x {pairs} pairs per launch",
            "hbm_frac_algorithmic": BYTES_PER_PARTICLE[w.dim] * w.n * steps_per_s / 1e9 / PEAK_HBM_GBS,
        },
    }
    if not args.no_cpu:
        cpu_steps = args.cpu_steps or {"c1": 100, "c2": 10, "c3": 3}[args.config]
        host = host_cpus()
        t_main = cpu_threads(args.cpu_threads, host)
        cb = cpu_baseline(w, cpu_steps, t_main, host)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
        # every CPU the process may use (affinity and cgroup quota), when that
        # differs from the main sample's count; a box's one-GPU share is
        # CPU_SHARE threads, so more are not started
        t_all = min(host["available"], CPU_SHARE)
        if t_all != t_main and args.config in ("c1", "c2"):
            out["cpu_all_available"] = cpu_baseline(w, cpu_steps, t_all, host)
        # the same restatement on one thread (the reference's LoadDistributor
        # with a single worker), one step; a 4M-particle clustered step would
        # take minutes on one thread, so C3 has none
        if args.config in ("c1", "c2"):
            out["cpu_single_thread"] = cpu_baseline(w, 1, 1, host)
    eng.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
