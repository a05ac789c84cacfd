"""Shared helpers: run the device step and the FP64 oracle on the same state.

Tolerances (unit box, the reference's own units):
  * cell keys, cell ranges (end offsets), permutation, particle order,
    per-particle neighbour count and neighbour-id hash, pair count: bit-exact;
  * velocities, and the FP32 accelerations the device exports in debug mode,
    against the oracle's a = (v_new - friction^(60 dt) v_old) / dt:
    |gpu - oracle| <= 1e-7 + 1e-5 |oracle|;
  * the acceleration derived from the device's stored FP32 velocities (the
    check when debug mode is off) also allows the storage rounding of v_new,
    |v_new| 2^-24 / dt;
  * positions (minimum image, so modulo 1 in a periodic world):
    |gpu - oracle| <= 2^-32 + dt (1e-7 + 1e-5 |v_oracle|)  -- the device rounds
    positions to its 2^-32 fixed-point grid.
"""
from __future__ import annotations

import numpy as np

import oracle

VEL_ABS = 1e-7
VEL_REL = 1e-5


def oracle_settings(w) -> "oracle.Settings":
    return oracle.Settings(matrix=w.matrix, rmax=w.rmax, friction=w.friction, dt=w.dt,
                           beta=w.beta, force=w.force, wrap=w.wrap, dim=w.dim)


def min_image(d: np.ndarray, wrap: bool) -> np.ndarray:
    d = np.abs(d)
    return np.minimum(d, 1.0 - d) if wrap else d


def check_one_step(eng, w, nthreads: int = 8, debug: bool = True) -> dict:
    """Step the device once from its current state and compare with the oracle
    stepping the exact same (downloaded) state.  Returns error statistics."""
    cur = eng.state()
    S = oracle_settings(w)
    ref = oracle.step(S, cur["pos"], cur["vel"], cur["types"], ids=cur["ids"],
                      nthreads=nthreads, debug=True)
    eng.set_debug(debug)
    eng.step(1)
    new = eng.state()
    cells = eng.cells()
    pairs = eng.pair_count()

    np.testing.assert_array_equal(cells["keys"], ref["keys"], err_msg="cell keys")
    np.testing.assert_array_equal(cells["cell_end"], ref["cell_end"], err_msg="cell ranges")
    np.testing.assert_array_equal(cells["perm"], ref["perm"], err_msg="permutation")
    np.testing.assert_array_equal(new["ids"], ref["ids"], err_msg="particle order")
    np.testing.assert_array_equal(new["types"], ref["types"], err_msg="types")
    assert pairs == ref["pair_count"], (pairs, ref["pair_count"])
    if debug:
        cnt, h = eng.neighbor_summary()
        np.testing.assert_array_equal(cnt, ref["nbr_count"], err_msg="neighbour counts")
        np.testing.assert_array_equal(h, ref["nbr_hash"], err_msg="neighbour sets")

    v_ref, v_gpu = ref["vel"], new["vel"]
    dv = np.abs(v_gpu - v_ref)
    bad = dv > VEL_ABS + VEL_REL * np.abs(v_ref)
    assert not bad.any(), f"velocity: {bad.sum()} components out of tolerance, max {dv.max():.3e}"

    ff = oracle.friction_factor(w.friction, w.dt)
    v_old = cur["vel"][cells["perm"]]
    a_ref = (v_ref - ff * v_old) / w.dt
    if debug:
        a_gpu = eng.accelerations().astype(np.float64)
        da = np.abs(a_gpu - a_ref)
        bad = da > VEL_ABS + VEL_REL * np.abs(a_ref)
        assert not bad.any(), f"acceleration: {bad.sum()} out of tolerance, max {da.max():.3e}"
    else:
        a_gpu = (v_gpu - ff * v_old) / w.dt
        da = np.abs(a_gpu - a_ref)
        bad = da > VEL_ABS + VEL_REL * np.abs(a_ref) + np.abs(v_gpu) * 2.0 ** -24 / w.dt
        assert not bad.any(), (f"acceleration (from stored velocities): {bad.sum()} out of "
                               f"tolerance, max {da.max():.3e}")

    dp = min_image(new["pos"] - ref["pos"], w.wrap)
    tol = 2.0 ** -32 + w.dt * (VEL_ABS + VEL_REL * np.abs(v_ref))
    bad = dp > tol
    assert not bad.any(), f"position: {bad.sum()} out of tolerance, max {dp.max():.3e}"
    return {"pairs": pairs, "max_dv": float(dv.max(initial=0)), "max_da": float(da.max(initial=0)),
            "max_dx": float(dp.max(initial=0))}


def free_run_error(eng, w, steps: int, nthreads: int = 8) -> float:
    """Max minimum-image position error after `steps` free-running steps of the
    device and of the oracle from the same initial state, matched by id."""
    eng.load_world(w)
    cur = eng.state()
    eng.step(steps)
    g = eng.state()
    S = oracle_settings(w)
    pos, vel, typ, ids = cur["pos"], cur["vel"], cur["types"], cur["ids"]
    for _ in range(steps):
        r = oracle.step(S, pos, vel, typ, ids=ids, nthreads=nthreads)
        pos, vel, typ, ids = r["pos"], r["vel"], r["types"], r["ids"]
    og = np.argsort(g["ids"])
    orf = np.argsort(ids)
    return float(min_image(g["pos"][og] - pos[orf], w.wrap).max(initial=0))
