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
