"""Closed forms of the normalized scores of a distance metric, vectorized over
(exact distance d, maximum M) arrays.

They restate the chain of CachedNormalizedMetricBase (details/distance.hpp:240-262)
for a metric whose _distance(s2, c) returns d if d <= c else c + 1
(details/distance.hpp:298-317), with NormSim_to_NormDist from
details/common.hpp:50-53.  numpy's float64 operations round like the C++
ones compiled without FMA contraction, so the results are bit patterns to
compare exactly.  tests/test_norm_closed_form.py pins them against the oracle
and the compiled reference.
"""
import numpy as np


def capped(d, c):
    """_distance(s2, c): d if d <= c else c + 1 (uint64)."""
    d = np.asarray(d, dtype=np.uint64)
    c = np.asarray(c, dtype=np.uint64)
    return np.where(d <= c, d, c + np.uint64(1)).astype(np.uint64)


def norm_distance(d, M, c):
    """_normalized_distance(s2, c) from the exact distance d and maximum M."""
    M = np.asarray(M, dtype=np.uint64)
    cutoff_distance = np.ceil(M.astype(np.float64) * c).astype(np.uint64)
    dist = capped(d, cutoff_distance).astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        norm = np.where(M != 0, dist / M.astype(np.float64), 0.0)
    return np.where(norm <= c, norm, 1.0)


def norm_sim_to_norm_dist(c):
    return min(1.0, 1.0 - c + 0.00001)


def norm_similarity(d, M, c):
    """_normalized_similarity(s2, c) from the exact distance d and maximum M."""
    nd = norm_distance(d, M, norm_sim_to_norm_dist(c))
    ns = 1.0 - nd
    return np.where(ns >= c, ns, 0.0)
