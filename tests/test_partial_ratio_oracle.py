"""fuzz::partial_ratio: pins of the oracle and of the window formulation the
kernels use.

* The reference's own assertions (test/tests-fuzz.cpp) are in
  tests/golden/known_answers.json (metric 10) and run in test_oracle_golden.
* The reference's cutoff checks of issues 206 and 210 (tests-fuzz.cpp:161-184):
  a cutoff just above the score gives 0, just below gives the score.
* The oracle restates the reference's search (bisection over the full windows
  with a pruning bound, running cutoffs, char-set filters, the swapped pass of
  equal lengths).  The kernels instead score every window and keep the best.
  `every_window` below is that formulation in plain Python, built on the
  oracle's ratio; it must equal the oracle (and the reference library, when
  built) on exhaustive tiny inputs and on random ones -- that is what licenses
  the kernels' reformulation.
"""
import itertools
import math

import numpy as np
import pytest

import _gen

PARTIAL, RATIO, SIM = 10, 6, 1


def every_window(oracle, s1, s2, cut):
    """max over every window of the reference's candidate set, each window with
    its own cutoff rule (fuzz_impl.hpp:81-175, 193-228)."""
    s1 = [int(x) for x in s1]
    s2 = [int(x) for x in s2]
    if len(s1) > len(s2):
        s1, s2 = s2, s1
    if cut > 100:
        return 0.0
    if not s1 or not s2:
        return 100.0 if len(s1) == len(s2) else 0.0

    def ratio(a, b):
        return float(oracle.score(RATIO, SIM, np.array(a, np.uint32), np.array(b, np.uint32), cutoff_d=cut))

    def lcs(a, b):
        return int(oracle.lib.orc_lcs_raw(np.array(a, np.uint32).ctypes.data_as(_u32p()), len(a),
                                          np.array(b, np.uint32).ctypes.data_as(_u32p()), len(b)))

    def windows(needle, hay):
        n1, n2 = len(needle), len(hay)
        best = 0.0
        if n2 > n1:
            d = min(2 * n1 - 2 * lcs(needle, hay[p:p + n1]) for p in range(n2 - n1))
            maximum = 2 * n1
            cutoff_dist = math.ceil(maximum * min(1.0, 1.0 - cut / 100 + 0.00001))
            if d < cutoff_dist:
                score = (1.0 - d / maximum) * 100
                if score >= cut:
                    best = max(best, score)
        cs = set(needle)
        for i in range(1, n1):
            if hay[i - 1] in cs:
                best = max(best, ratio(needle, hay[:i]))
        for i in range(n2 - n1, n2):
            if hay[i] in cs:
                best = max(best, ratio(needle, hay[i:]))
        return best

    best = windows(s1, s2)
    if len(s1) == len(s2):
        best = max(best, windows(s2, s1))
    return best


def _u32p():
    import ctypes
    return ctypes.POINTER(ctypes.c_uint32)


def test_every_window_exhaustive_tiny(oracle):
    # all strings over {a, b} up to length 4 against all up to length 5, three cutoffs
    alpha = [97, 98]
    small = [list(t) for L in range(0, 5) for t in itertools.product(alpha, repeat=L)]
    large = [list(t) for L in range(0, 6) for t in itertools.product(alpha, repeat=L)][::3]
    for cut in (0.0, 60.0, 85.0):
        got = oracle.cdist(PARTIAL, SIM, [np.array(s, np.uint32) for s in small],
                           [np.array(s, np.uint32) for s in large], cutoff_d=cut)
        for i, a in enumerate(small):
            for j, b in enumerate(large):
                exp = every_window(oracle, a, b, cut)
                assert got[i, j] == exp, (a, b, cut, got[i, j], exp)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_every_window_random(oracle, seed):
    rng = np.random.default_rng(seed)
    q = _gen.random_strings(rng, 12, 0, 24, alpha_n=3 + seed)
    c = _gen.random_strings(rng, 14, 0, 60, alpha_n=3 + seed)
    c += [np.concatenate([_gen.random_strings(rng, 1, 0, 20, alpha_n=3)[0], _gen.mutate(rng, q[i], 2),
                          _gen.random_strings(rng, 1, 0, 20, alpha_n=3)[0]]) for i in range(6)]
    for cut in (0.0, 50.0, 77.0, 100.0):
        got = oracle.cdist(PARTIAL, SIM, q, c, cutoff_d=cut)
        for i, a in enumerate(q):
            for j, b in enumerate(c):
                assert got[i, j] == every_window(oracle, a, b, cut), (i, j, cut)


def test_every_window_vs_reference(oracle, reference):
    rng = np.random.default_rng(11)
    q = _gen.random_strings(rng, 10, 1, 70, alpha_n=4)
    c = _gen.random_strings(rng, 10, 1, 90, alpha_n=4)
    for cut in (0.0, 70.0):
        exp = reference.cdist(PARTIAL, SIM, q, c, cutoff_d=cut)
        for i, a in enumerate(q):
            for j, b in enumerate(c):
                assert exp[i, j] == every_window(oracle, a, b, cut), (i, j, cut)


@pytest.mark.parametrize("s1,s2", [(b"South Korea", b"North Korea"), (b"bc", b"bca")])
def test_cutoff_boundary_issues(oracle, s1, s2):
    # tests-fuzz.cpp:161-184, both argument orders
    for a, b in ((s1, s2), (s2, s1)):
        score = float(oracle.score(PARTIAL, SIM, a, b, cutoff_d=0.0))
        assert score > 0
        assert float(oracle.score(PARTIAL, SIM, a, b, cutoff_d=score + 0.0001)) == 0.0
        assert abs(float(oracle.score(PARTIAL, SIM, a, b, cutoff_d=score - 0.0001)) - score) <= 1e-6


def test_each_window_rule_is_pinned(oracle):
    # Which known answer fails when a candidate rule is dropped (scores without the
    # rule in brackets): full windows -- tests-fuzz.cpp:157 (0) and :158 (84.6);
    # prefix windows -- :195 (49.5); suffix windows -- :77 (96), :207/:208 (97.7).
    # The swapped pass of equal lengths is pinned here: "aaa" vs "aba" scores 80
    # only through it (66.7 without).
    assert float(oracle.score(PARTIAL, SIM, b"aaa", b"aba")) == 80.0
    assert float(oracle.score(PARTIAL, SIM, b"aba", b"aaa")) == 80.0
    assert every_window(oracle, b"aaa", b"aba", 0.0) == 80.0


def test_swapped_pass_vs_reference(reference):
    assert float(reference.cdist(PARTIAL, SIM, [b"aaa"], [b"aba"])[0, 0]) == 80.0


def test_window_score_depends_on_the_fraction_only(oracle):
    """The kernels score one window per pair: the one with the largest 2 lcs / (l1 + L).
    That is exact when CachedRatio's score of a window is a non-decreasing function of
    x = 1 - d / M alone (d = l1 + L - 2 lcs, M = l1 + L), so that equal fractions score
    equal doubles and a smaller fraction never scores more.  Checked on every
    (l1, L, lcs) with l1, L <= 18 and cutoffs that put 1 - c/100 + 1e-5 on and beside
    such fractions."""
    cuts = [0.0, 25.0, 50.0, 60.0, 75.0, 80.0, 100.0 * (1 - 1 / 3) + 0.001,
            100.0 * (1 - 0.3) - 0.001, 100.0 * (1 - 0.25 + 0.00001), 100.0 * (1 - 7 / 15 + 0.00001)]
    cases = []
    for l1 in range(1, 19):
        for L in range(1, 19):
            for lcs in range(0, min(l1, L) + 1):
                cases.append((l1, L, lcs))
    queries = [np.array([97] * lcs + [98] * (l1 - lcs), np.uint32) for l1, L, lcs in cases]
    windows = [np.array([97] * lcs + [99] * (L - lcs), np.uint32) for l1, L, lcs in cases]
    from fractions import Fraction
    for cut in cuts:
        by_fraction = {}
        for (l1, L, lcs), a, b in zip(cases, queries, windows):
            v = float(oracle.score(RATIO, SIM, a, b, cutoff_d=cut))
            by_fraction.setdefault(Fraction(2 * lcs, l1 + L), set()).add(v)
        fracs = sorted(by_fraction)
        for f in fracs:
            assert len(by_fraction[f]) == 1, (cut, f, by_fraction[f])
        vals = [next(iter(by_fraction[f])) for f in fracs]
        assert all(x <= y for x, y in zip(vals, vals[1:])), cut
