"""The oracle's Damerau-Levenshtein (Zhao and Sahni's algorithm as
damerau_levenshtein_distance_zhao runs it, DamerauLevenshtein_impl.hpp:36-121,
through experimental::CachedDamerauLevenshtein, DamerauLevenshtein.hpp:112-140),
pinned three ways:

* against the compiled reference for every kind and cutoff and 8/16/32-bit units;
* against the reference's own known answers (tests-DamerauLevenshtein.cpp:104-133,
  in tests/golden/known_answers.json);
* against the Lowrance-Wagner recurrence written out below (the unrestricted
  edit distance with adjacent transpositions, a different formulation: a full
  matrix and the last-occurrence term D(k-1, l-1) + (i-k-1) + 1 + (j-l-1)), and
  against symmetry in its two arguments, on exhaustive tiny strings and random
  ones.  Dropping either of Zhao's two transposition cases, or swapping the FR /
  T roles, fails these.
"""
import itertools

import numpy as np
import pytest

import _gen

DL = 18


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def lowrance_wagner(a, b):
    inf = len(a) + len(b) + 1
    da = {}
    D = [[0] * (len(b) + 2) for _ in range(len(a) + 2)]
    D[0][0] = inf
    for i in range(len(a) + 1):
        D[i + 1][0] = inf
        D[i + 1][1] = i
    for j in range(len(b) + 1):
        D[0][j + 1] = inf
        D[1][j + 1] = j
    for i in range(1, len(a) + 1):
        db = 0
        for j in range(1, len(b) + 1):
            k, l = da.get(b[j - 1], 0), db
            cost = 0 if a[i - 1] == b[j - 1] else 1
            if cost == 0:
                db = j
            D[i + 1][j + 1] = min(D[i][j] + cost, D[i + 1][j] + 1, D[i][j + 1] + 1,
                                  D[k][l] + (i - k - 1) + 1 + (j - l - 1))
        da[a[i - 1]] = i
    return D[len(a) + 1][len(b) + 1]


def test_exhaustive_tiny_vs_lowrance_wagner(oracle):
    words = [()] + [w for n in range(1, 5) for w in itertools.product((97, 98, 99), repeat=n)]
    q = [np.array(w, dtype=np.uint32) for w in words]
    got = oracle.cdist(DL, 0, q, q)
    for x, a in enumerate(words):
        for y, b in enumerate(words):
            assert int(got[x, y]) == lowrance_wagner(a, b), f"{a} {b}"


def test_random_vs_lowrance_wagner_and_symmetry(oracle):
    rng = np.random.default_rng(3)
    q = _gen.random_strings(rng, 40, 0, 30, alpha_n=4) + _gen.random_strings(rng, 5, 0, 60, alpha_n=2)
    c = _gen.random_strings(rng, 40, 0, 40, alpha_n=4)
    d = oracle.cdist(DL, 0, q, c)
    dt = oracle.cdist(DL, 0, c, q)
    assert np.array_equal(d, dt.T)
    for i in range(0, len(q), 2):
        for j in range(0, len(c), 3):
            assert int(d[i, j]) == lowrance_wagner(tuple(int(x) for x in q[i]), tuple(int(x) for x in c[j]))


def _cutoffs(kind):
    if kind in (0, 1):
        return [(None, None), (0, None), (1, None), (2, None), (5, None), (17, None)]
    return [(None, None), (None, 0.0), (None, 0.3), (None, 0.7), (None, 0.75), (None, 0.8), (None, 1.0)]


@pytest.mark.parametrize("alpha", [2, 4, 26])
def test_vs_reference(oracle, reference, alpha):
    rng = np.random.default_rng(50 + alpha)
    q = _gen.random_strings(rng, 24, 0, 90, alpha_n=alpha)
    c = _gen.random_strings(rng, 40, 0, 140, alpha_n=alpha)
    c += [_gen.mutate(rng, q[i], int(rng.integers(0, 5)), alpha_n=alpha) for i in range(len(q))]
    # adjacent swaps: the transposition terms
    for s in q[:12]:
        t = s.copy()
        for p in range(0, len(t) - 1, 5):
            t[p], t[p + 1] = t[p + 1], t[p]
        c.append(t)
    q += _gen.edge_strings()[:10]
    for kind in (0, 1, 2, 3):
        for cu, cd in _cutoffs(kind):
            exp = reference.cdist(DL, kind, q, c, cutoff_u=cu, cutoff_d=cd)
            got = oracle.cdist(DL, kind, q, c, cutoff_u=cu, cutoff_d=cd)
            assert _same(got, exp), f"alpha {alpha} kind {kind} cutoff {cu}/{cd}"


@pytest.mark.parametrize("width", [2, 4])
def test_vs_reference_wide_units(oracle, reference, width):
    rng = np.random.default_rng(60 + width)
    lo = 0x4E00 if width == 2 else 0x1F300
    q = _gen.random_strings(rng, 20, 0, 70, alpha_lo=lo, alpha_n=6)
    c = _gen.random_strings(rng, 30, 0, 90, alpha_lo=lo, alpha_n=6)
    for kind in (0, 3):
        exp = reference.cdist(DL, kind, q, c, char_width=width)
        got = oracle.cdist(DL, kind, q, c)
        assert _same(got, exp)
