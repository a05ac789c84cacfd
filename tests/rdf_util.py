"""Type-pair radial distribution of a particle-life state, and a two-sample
statistic for comparing two such distributions.

Long runs of the step are chaotic (every step feeds back the last;
backend/Physics.java:397-450), so two correct implementations drift apart
particle by particle while the structure they form stays alike.  A long run
is therefore compared by the distribution of pair distances per unordered
type pair, histogrammed up to a given radius.
"""
from __future__ import annotations

import numpy as np
from scipy.spatial import cKDTree


def type_pair_rdf(pos: np.ndarray, types: np.ndarray, ntypes: int, r: float, nbins: int,
                  dim: int = 2, wrap: bool = True) -> np.ndarray:
    """Counts of unordered pairs (i < j) with 0 < |d| <= r (minimum image when
    wrap), by (type pair, distance bin): shape (ntypes * (ntypes + 1) / 2, nbins)."""
    p = np.ascontiguousarray(pos[:, :dim], dtype=np.float64)
    if wrap:
        p = p - np.floor(p)             # the tree's periodic box is [0, 1)
        p[p >= 1.0] = 0.0
    tree = cKDTree(p, boxsize=1.0 if wrap else None)
    pairs = tree.query_pairs(r, output_type="ndarray")
    if len(pairs) == 0:
        return np.zeros((ntypes * (ntypes + 1) // 2, nbins), np.int64)
    i, j = pairs[:, 0], pairs[:, 1]
    d = p[j] - p[i]
    if wrap:
        d -= np.round(d)
    dist = np.sqrt((d * d).sum(axis=1))
    keep = dist > 0
    a = np.minimum(types[i], types[j])[keep]
    b = np.maximum(types[i], types[j])[keep]
    pair_index = a * ntypes - a * (a - 1) // 2 + (b - a)       # row-major upper triangle
    bins = np.minimum((dist[keep] / r * nbins).astype(np.int64), nbins - 1)
    out = np.zeros((ntypes * (ntypes + 1) // 2, nbins), np.int64)
    np.add.at(out, (pair_index, bins), 1)
    return out


def two_sample_chi2(h1: np.ndarray, h2: np.ndarray) -> tuple[float, int]:
    """sum (h1 - h2)^2 / (h1 + h2) over the non-empty cells, and their number.
    For two independent Poisson samples of one distribution its expectation is
    about the number of non-empty cells; identical histograms give 0."""
    h1 = h1.astype(np.float64).ravel()
    h2 = h2.astype(np.float64).ravel()
    m = (h1 + h2) > 0
    return float(((h1[m] - h2[m]) ** 2 / (h1[m] + h2[m])).sum()), int(m.sum())
