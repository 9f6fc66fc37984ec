"""Pins the closed-form normalization of tests/_norm.py (used by the exhaustive GPU
check over every (distance, maximum <= 8192)) against the oracle and the compiled
reference, on Hamming distances of 'a'^i against 'a'^j (d = |i - j|, maximum =
max(i, j), CachedHamming with pad = true, Hamming_impl.hpp:16-35)."""
import numpy as np
import pytest

import _norm

N = 300
NORM_DIST_CUTOFFS = [1.0, 0.0, 0.1, 0.2, 1.0 / 3.0, 0.5, 0.75]
NORM_SIM_CUTOFFS = [0.0, 0.2, 0.5, 2.0 / 3.0, 0.8, 0.9, 1.0]


def _grid(n):
    strs = [np.full(i, 97, dtype=np.uint32) for i in range(n + 1)]
    i = np.arange(n + 1, dtype=np.uint64)
    d = np.abs(i[:, None].astype(np.int64) - i[None, :].astype(np.int64)).astype(np.uint64)
    M = np.maximum(i[:, None], i[None, :])
    return strs, d, M


def _same(a, b):
    return np.array_equal(np.asarray(a, np.float64).view(np.uint64), np.asarray(b, np.float64).view(np.uint64))


@pytest.mark.parametrize("c", NORM_DIST_CUTOFFS)
def test_norm_distance_vs_oracle(oracle, c):
    strs, d, M = _grid(N)
    exp = oracle.cdist(7, 2, strs, strs, cutoff_d=c)
    assert _same(_norm.norm_distance(d, M, c), exp)


@pytest.mark.parametrize("c", NORM_SIM_CUTOFFS)
def test_norm_similarity_vs_oracle(oracle, c):
    strs, d, M = _grid(N)
    exp = oracle.cdist(7, 3, strs, strs, cutoff_d=c)
    assert _same(_norm.norm_similarity(d, M, c), exp)


@pytest.mark.parametrize("c", [0.2, 0.8, 2.0 / 3.0])
def test_closed_form_vs_reference(reference, c):
    strs, d, M = _grid(120)
    assert _same(_norm.norm_distance(d, M, c), reference.cdist(7, 2, strs, strs, cutoff_d=c))
    assert _same(_norm.norm_similarity(d, M, c), reference.cdist(7, 3, strs, strs, cutoff_d=c))


def test_grid_catches_plausible_mistakes(oracle):
    """Each of these slips in the chain changes some value on the grid, so the grid pins it."""
    strs, d, M = _grid(N)
    exp_sim = oracle.cdist(7, 3, strs, strs, cutoff_d=0.8)
    exp_dist = oracle.cdist(7, 2, strs, strs, cutoff_d=0.2)

    def sim_without_widening(c):  # NormSim_to_NormDist without the +1e-5
        nd = _norm.norm_distance(d, M, min(1.0, 1.0 - c))
        ns = 1.0 - nd
        return np.where(ns >= c, ns, 0.0)

    def dist_strict_cut(c):  # norm < c instead of norm <= c
        cd = np.ceil(M.astype(np.float64) * c).astype(np.uint64)
        dist = _norm.capped(d, cd).astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            norm = np.where(M != 0, dist / M.astype(np.float64), 0.0)
        return np.where(norm < c, norm, 1.0)

    def dist_empty_as_one(c):  # maximum 0 giving 1.0 instead of 0.0
        v = _norm.norm_distance(d, M, c)
        return np.where(M == 0, 1.0, v)

    assert not _same(sim_without_widening(0.8), exp_sim)
    assert not _same(dist_strict_cut(0.2), exp_dist)
    assert not _same(dist_empty_as_one(0.2), exp_dist)


@pytest.mark.parametrize("metric", [0, 1, 3, 6])
@pytest.mark.parametrize("kind,c", [(2, 0.2), (2, 1.0 / 3.0), (3, 0.8), (3, 2.0 / 3.0), (3, 0.0)])
def test_levenshtein_family_closed_form(oracle, metric, kind, c):
    """The same closed forms for 'a'^i against 'a'^j under Levenshtein / OSA (d = |i - j|, maximum
    max(i, j)), Indel (d = |i - j|, maximum i + j) and ratio (Indel normalized similarity at c / 100,
    times 100): the expected values of the GPU's normalized epilogue."""
    n = 120
    strs, d, M = _grid(n)
    if metric in (1, 6):
        i = np.arange(n + 1, dtype=np.uint64)
        M = i[:, None] + i[None, :]
    if metric == 6:
        if kind != 3:
            pytest.skip("ratio is a similarity")
        exp = oracle.cdist(6, 1, strs, strs, cutoff_d=c * 100)
        assert _same(_norm.norm_similarity(d, M, c) * 100, exp)
        return
    exp = oracle.cdist(metric, kind, strs, strs, cutoff_d=c)
    f = _norm.norm_distance if kind == 2 else _norm.norm_similarity
    assert _same(f(d, M, c), exp)
