"""Long runs compared statistically: the device and the oracle each run a
world for many steps from the same initial state; the oracle's own spread is
measured by re-running it from states nudged by one unit of the device's
2^-32 position grid.  Shared by the GPU test and its CPU pinning test."""
from __future__ import annotations

import numpy as np

import oracle
from parity_util import min_image, oracle_settings
from rdf_util import two_sample_chi2, type_pair_rdf

RDF_BINS = 16


def oracle_run(w, pos, steps: int, nthreads: int = 8, settings=None):
    """steps oracle steps from (pos, w.vel, w.types); returns (pos, types, ids)."""
    S = settings or oracle_settings(w)
    vel, typ, ids = w.vel, w.types, None
    for _ in range(steps):
        r = oracle.step(S, pos, vel, typ, ids=ids, nthreads=nthreads)
        pos, vel, typ, ids = r["pos"], r["vel"], r["types"], r["ids"]
    return pos, typ, ids


def nudged(w, k: int) -> np.ndarray:
    """w.pos with every coordinate moved by +-2^-32 (seed k), wrapped."""
    rng = np.random.default_rng(1000 + k)
    p = w.pos.copy()
    p[:, :w.dim] += rng.choice([-1.0, 1.0], size=(w.n, w.dim)) * 2.0 ** -32
    return p - np.floor(p) if w.wrap else np.clip(p, 0.0, 1.0)


def rdf(w, pos, types) -> np.ndarray:
    return type_pair_rdf(pos, types, w.ntypes, 2.0 * w.rmax, RDF_BINS, dim=w.dim, wrap=w.wrap)


def max_error(w, a, b) -> float:
    """Largest minimum-image position difference of two runs, matched by id."""
    pa, _, ia = a
    pb, _, ib = b
    return float(min_image(pa[np.argsort(ia)] - pb[np.argsort(ib)], w.wrap).max(initial=0))


def spread(w, steps: int, nruns: int = 3, nthreads: int = 8):
    """The oracle's reference run and its spread under 2^-32 nudges: the
    largest RDF statistic and position error of a nudged run against it."""
    ref = oracle_run(w, w.pos, steps, nthreads)
    h = rdf(w, ref[0], ref[1])
    chi, err = 0.0, 0.0
    for k in range(nruns):
        r = oracle_run(w, nudged(w, k), steps, nthreads)
        chi = max(chi, two_sample_chi2(h, rdf(w, r[0], r[1]))[0])
        err = max(err, max_error(w, ref, r))
    return ref, h, chi, err


def chi2_tolerance(spread_chi2: float) -> float:
    """Allowed RDF statistic against the oracle: ten times the oracle's own
    spread, plus 2 -- one pair moving to a neighbouring distance bin adds at
    most 2."""
    return 2.0 + 10.0 * spread_chi2


def error_tolerance(spread_err: float) -> float:
    """Allowed position error against the oracle: ten times its own spread."""
    return 10.0 * spread_err
