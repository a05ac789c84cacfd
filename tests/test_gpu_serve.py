"""The continuous-batching server: jobs that join and leave a running batch
end bit-identical to the same jobs run alone, each after exactly its steps."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trace():
    from particle_life import worlds

    jobs = []
    specs = [  # (n, dim, ntypes, rmax, steps, arrival)
        (3000, 2, 5, 0.02, 40, 0.00), (500, 3, 4, 0.05, 9, 0.00), (7000, 2, 8, 0.015, 23, 0.01),
        (1, 2, 4, 0.03, 5, 0.02), (2500, 3, 6, 0.04, 31, 0.02), (4000, 2, 10, 0.03, 17, 0.05),
        (1200, 3, 7, 0.06, 12, 0.08), (6000, 2, 4, 0.025, 50, 0.10),
    ]
    for j, (n, dim, t, rmax, steps, arr) in enumerate(specs):
        w = worlds.uniform_world(300 + j, n, t, rmax=rmax, dt=0.01, beta=0.3, dim=dim)
        jobs.append(dict(job=j, arrival=arr, steps=steps, world=w))
    return jobs


def test_served_jobs_match_alone():
    from particle_life import Engine
    from particle_life.serve import serve

    print("\n[serve] case=served_vs_alone", flush=True)
    jobs = _trace()
    st = serve(jobs, quantum=8, keep_state=True)
    assert len(st.results) == len(jobs)
    for r, j in zip(st.results, jobs):
        assert r.job == j["job"] and r.done >= r.admitted >= 0
        e = Engine(0)
        e.load_world(j["world"])
        e.step(j["steps"])
        alone = e.state()
        e.close()
        for k in ("pos", "vel", "ids", "types"):
            np.testing.assert_array_equal(r.state[k], alone[k], err_msg=f"job {r.job}: {k}")
    s = st.summary()
    assert s["jobs"] == len(jobs) and s["p99_s"] >= s["p50_s"] > 0


def test_tight_budget_serialises_admission():
    """With a budget below two jobs' footprint the server still finishes every
    job (one at a time) with the same results."""
    from particle_life.serve import serve

    print("\n[serve] case=tight_budget", flush=True)
    jobs = _trace()[:4]
    a = serve(jobs, quantum=16, keep_state=True)
    b = serve(jobs, quantum=16, keep_state=True, budget_bytes=1.0)
    for ra, rb in zip(a.results, b.results):
        for k in ("pos", "vel", "ids"):
            np.testing.assert_array_equal(ra.state[k], rb.state[k])


# the first 22 arrivals of the served workload (among them job 5, its largest
# 2D world) and job 163, its largest 3D world, which is moved up to arrive
# while the others run
C5_JOBS = list(range(22)) + [163]
C5_ORACLE_JOBS = (5, 163, 3, 19, 11)


def test_served_workload_at_full_size():
    """Jobs of the served workload (worlds.config5_jobs) at their real sizes,
    dimensions and step counts -- 2D and 3D worlds of 2k-262k particles run
    for 130-1991 steps, arriving at their trace times -- each end
    bit-identical to the same job run alone; the final states of five of them
    (the largest 2D and 3D worlds among them) then step once against the
    oracle."""
    from particle_life import Engine, worlds
    from particle_life.serve import serve
    from parity_util import check_one_step

    print("\n[serve] case=c5_full_size", flush=True)
    trace = worlds.config5_jobs(200)
    jobs = []
    for k in C5_JOBS:
        j = trace[k]
        jobs.append(dict(job=j["job"], arrival=min(j["arrival"], 0.9), steps=j["steps"],
                         world=worlds.job_world(j)))
    assert max(j["world"].n for j in jobs if j["world"].dim == 2) == 260_132
    assert max(j["world"].n for j in jobs if j["world"].dim == 3) == 258_733
    st = serve(jobs, quantum=32, keep_state=True)
    s = st.summary()
    print(f"  served {s['jobs']} jobs: p50 {s['p50_s']:.2f} s, p99 {s['p99_s']:.2f} s, "
          f"{s['particle_steps_per_s']:.3g} particle-steps/s", flush=True)
    assert len(st.results) == len(jobs)
    by_job = {j["job"]: j for j in jobs}
    for r in st.results:
        j = by_job[r.job]
        with Engine(0) as e:
            e.load_world(j["world"])
            e.step(j["steps"])
            alone = e.state()
        for key in ("pos", "vel", "ids", "types"):
            np.testing.assert_array_equal(r.state[key], alone[key], err_msg=f"job {r.job}: {key}")
    for r in st.results:
        if r.job not in C5_ORACLE_JOBS:
            continue
        w = by_job[r.job]["world"]
        print(f"  oracle check of job {r.job} (n={w.n}, dim={w.dim}) after {r.steps} steps",
              flush=True)
        with Engine(0) as e:
            e.configure_world(w)
            e.set_state(r.state["pos"], r.state["vel"], r.state["types"])
            st1 = check_one_step(e, w, nthreads=16, debug=True)
        print(f"    {st1}", flush=True)
