"""The oracle's Levenshtein with general weights (generalized Wagner-Fischer,
Levenshtein_impl.hpp:85-155, reached from CachedLevenshtein::_distance,
Levenshtein.hpp:445-477), pinned three ways:

* against the compiled reference (oracle/_ref), every kind and cutoff, for weights
  that take the general path (insert != delete, or a replace weight strictly between
  the common weight and twice it, or below it);
* against an enumeration of every alignment of tiny strings (no recurrence: each
  alignment's cost is summed directly), which catches a swapped insert / delete
  weight, a dropped term or a wrong diagonal;
* through identities the weighted distance obeys (empty strings, swapping the
  strings with the insert / delete weights, clamping the replace weight at
  insert + delete).
"""
import functools
import itertools

import numpy as np
import pytest

import _gen

GENERAL = [(1, 2, 3), (2, 1, 3), (1, 1, 0), (2, 2, 3), (2, 2, 1), (3, 5, 4), (1, 3, 10), (5, 1, 1), (0, 1, 1),
           (1, 0, 2), (4, 4, 5), (7, 3, 2)]


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def _cutoffs(kind):
    if kind in (0, 1):
        return [(None, None), (0, None), (1, None), (4, None), (11, None), (40, None)]
    return [(None, None), (None, 0.0), (None, 0.3), (None, 0.5), (None, 0.75), (None, 1.0)]


@pytest.mark.parametrize("weights", GENERAL)
def test_general_weights_vs_reference(oracle, reference, weights):
    rng = np.random.default_rng(sum(weights) * 31 + weights[0])
    q = _gen.random_strings(rng, 18, 0, 70, alpha_n=4) + _gen.random_strings(rng, 4, 100, 300, alpha_n=3)
    c = _gen.random_strings(rng, 25, 0, 120, alpha_n=4)
    c += [_gen.mutate(rng, q[i], int(rng.integers(0, 6))) for i in range(len(q))]
    for kind in (0, 1, 2, 3):
        for cu, cd in _cutoffs(kind):
            exp = reference.cdist(0, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd)
            got = oracle.cdist(0, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd)
            assert _same(got, exp), f"weights {weights} kind {kind} cutoff {cu}/{cd}"


def _alignment_min(a, b, ins, dele, rep):
    """Cheapest alignment of a onto b by explicit enumeration of edit paths (memoized on
    the remaining suffixes only to keep the enumeration finite in time)."""

    @functools.lru_cache(maxsize=None)
    def best(i, j):
        if i == len(a) and j == len(b):
            return 0
        options = []
        if i < len(a):
            options.append(dele + best(i + 1, j))  # drop a[i]
        if j < len(b):
            options.append(ins + best(i, j + 1))  # insert b[j]
        if i < len(a) and j < len(b):
            options.append((0 if a[i] == b[j] else rep) + best(i + 1, j + 1))
        return min(options)

    return best(0, 0)


def _all_paths_min(a, b, ins, dele, rep):
    """Every edit path as a sequence of moves, summed without any recurrence."""
    best = None
    n_moves_max = len(a) + len(b)
    for n_moves in range(max(len(a), len(b)), n_moves_max + 1):
        for moves in itertools.product("DIS", repeat=n_moves):
            i = j = cost = 0
            ok = True
            for mv in moves:
                if mv == "D":
                    if i >= len(a):
                        ok = False
                        break
                    cost += dele
                    i += 1
                elif mv == "I":
                    if j >= len(b):
                        ok = False
                        break
                    cost += ins
                    j += 1
                else:
                    if i >= len(a) or j >= len(b):
                        ok = False
                        break
                    cost += 0 if a[i] == b[j] else rep
                    i += 1
                    j += 1
            if ok and i == len(a) and j == len(b):
                best = cost if best is None else min(best, cost)
    return best


@pytest.mark.parametrize("weights", GENERAL)
def test_general_weights_vs_enumeration(oracle, weights):
    ins, dele, rep = weights
    words = [()] + [w for n in range(1, 4) for w in itertools.product((97, 98, 99), repeat=n)]
    rng = np.random.default_rng(5)
    pick = [words[i] for i in rng.choice(len(words), size=16, replace=False)] + [(), (97, 98, 99), (99, 99)]
    q = [np.array(w, dtype=np.uint32) for w in pick]
    got = oracle.cdist(0, 0, q, q, weights=weights)
    for x, a in enumerate(pick):
        for y, b in enumerate(pick):
            exp = _all_paths_min(a, b, ins, dele, rep)
            assert int(got[x, y]) == exp, f"{a} -> {b} weights {weights}: oracle {got[x, y]} paths {exp}"


@pytest.mark.parametrize("weights", GENERAL)
def test_general_weights_identities(oracle, weights):
    ins, dele, rep = weights
    rng = np.random.default_rng(17 + ins)
    q = _gen.random_strings(rng, 12, 0, 40, alpha_n=3) + [np.zeros(0, np.uint32)]
    c = _gen.random_strings(rng, 15, 0, 50, alpha_n=3) + [np.zeros(0, np.uint32)]
    d = oracle.cdist(0, 0, q, c, weights=weights).astype(np.int64)
    # empty strings: every unit deleted / inserted
    assert all(d[i, len(c) - 1] == len(q[i]) * dele for i in range(len(q)))
    assert all(d[len(q) - 1, j] == len(c[j]) * ins for j in range(len(c)))
    # swapping the strings swaps the insert and delete weights
    dt = oracle.cdist(0, 0, c, q, weights=(dele, ins, rep)).astype(np.int64)
    assert np.array_equal(d, dt.T)
    # a replace weight above insert + delete is never used
    if rep > ins + dele:
        dc = oracle.cdist(0, 0, q, c, weights=(ins, dele, ins + dele)).astype(np.int64)
        assert np.array_equal(d, dc)
    # the recurrence against the memoized alignment search on a few longer pairs
    for i in range(0, len(q), 3):
        for j in range(0, len(c), 4):
            a, b = tuple(int(x) for x in q[i]), tuple(int(x) for x in c[j])
            assert d[i, j] == _alignment_min(a, b, ins, dele, rep)
