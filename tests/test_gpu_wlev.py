"""GPU parity of Levenshtein with general weights (generalized Wagner-Fischer,
Levenshtein_impl.hpp:85-155) against the oracle, bit for bit: the tile kernel
(queries of at most 64 units), the long kernel (one wave per pair, the query's
rows across the lanes), its swapped form (queries over 1024 units against
choices of at most 1024) and its strips (both strings over 1024 units), every
kind and cutoff, per-query cutoffs and top-k."""
import numpy as np
import pytest

import _gen
from _topk import expected_topk

pytestmark = pytest.mark.gpu

GENERAL = [(1, 2, 3), (2, 1, 3), (1, 1, 0), (2, 2, 3), (2, 2, 1), (3, 5, 4), (1, 3, 10), (5, 1, 1), (0, 1, 1),
           (1, 0, 2), (4, 4, 5)]


def _cutoffs(kind):
    if kind in (0, 1):
        return [(None, None), (0, None), (1, None), (4, None), (11, None), (40, None), (300, None)]
    return [(None, None), (None, 0.0), (None, 0.3), (None, 0.5), (None, 0.75), (None, 0.9), (None, 1.0)]


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def check(ctx, oracle, queries, choices, weights, width=1, kinds=(0, 1, 2, 3)):
    import rapidfuzz_amd as rfa

    print(f"weights {weights} width {width} nq {len(queries)} nc {len(choices)}", flush=True)
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=width), weights=weights)
    cs = ctx.strings(choices, width=width)
    for kind in kinds:
        for cu, cd in _cutoffs(kind):
            got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd)
            exp = oracle.cdist(0, kind, queries, choices, weights=weights, cutoff_u=cu, cutoff_d=cd)
            if not _same(got, exp):
                bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
                i, j = bad[0]
                raise AssertionError(f"weights {weights} kind {kind} cutoff {cu}/{cd}: {len(bad)} differ; q{i} "
                                     f"(len {len(queries[i])}) c{j} (len {len(choices[j])}): got {got[i, j]!r} "
                                     f"expected {exp[i, j]!r}")
    sc.close()


@pytest.mark.parametrize("weights", GENERAL)
def test_tile_path(ctx, oracle, weights):
    rng = np.random.default_rng(200 + 7 * weights[0] + weights[2])
    queries = _gen.random_strings(rng, 90, 0, 64, alpha_n=4) + [np.zeros(0, np.uint32)]
    choices = _gen.random_strings(rng, 60, 0, 150, alpha_n=4) + [np.zeros(0, np.uint32)]
    choices += [_gen.mutate(rng, queries[i], int(rng.integers(0, 6))) for i in range(30)]
    check(ctx, oracle, queries, choices, weights)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("weights", [(1, 2, 3), (3, 1, 2)])
def test_tile_path_wide_units(ctx, oracle, width, weights):
    rng = np.random.default_rng(300 + width)
    lo = 0x4E00 if width == 2 else 0x1F300
    queries = _gen.random_strings(rng, 70, 0, 64, alpha_lo=lo, alpha_n=40)
    choices = _gen.random_strings(rng, 50, 0, 100, alpha_lo=lo, alpha_n=40)
    choices += [_gen.mutate(rng, queries[i], 3, alpha_lo=lo, alpha_n=40) for i in range(20)]
    check(ctx, oracle, queries, choices, weights, width=width)


@pytest.mark.parametrize("weights", [(1, 2, 3), (2, 2, 1), (1, 3, 10), (5, 1, 1)])
def test_long_path(ctx, oracle, weights):
    rng = np.random.default_rng(400 + weights[1])
    # 65..1024 units: 2 .. 16 rows per lane
    queries = [rng.integers(97, 101, size=int(n)).astype(np.uint32) for n in (65, 100, 128, 129, 200, 300, 513,
                                                                                  700, 1024)]
    choices = _gen.random_strings(rng, 20, 0, 400, alpha_n=4) + [np.zeros(0, np.uint32)]
    choices += [_gen.mutate(rng, q, int(rng.integers(0, 20))) for q in queries]
    check(ctx, oracle, queries, choices, weights)


@pytest.mark.parametrize("weights", [(1, 2, 3), (4, 1, 3)])
def test_long_path_swapped(ctx, oracle, weights):
    """Queries over 1024 units hold the choice across the lanes with insert and delete swapped."""
    rng = np.random.default_rng(500 + weights[0])
    queries = [rng.integers(97, 100, size=n).astype(np.uint32) for n in (1025, 1300)]
    queries.append(rng.integers(97, 100, size=80).astype(np.uint32))  # a short one beside them
    choices = _gen.random_strings(rng, 12, 0, 1024, alpha_n=3) + [np.zeros(0, np.uint32)]
    choices += [queries[0][:1024], _gen.mutate(rng, queries[1][:1000], 10, alpha_n=3)]
    check(ctx, oracle, queries, choices, weights)


@pytest.mark.parametrize("weights", [(1, 2, 3), (3, 1, 2), (2, 2, 1), (1, 1, 5)])
def test_long_path_strips(ctx, oracle, weights):
    """Both strings over 1024 units: the query in strips of 1024 rows (2 and 3 strips, a strip
    boundary exactly at the end), against choices on either side of 1024 units."""
    rng = np.random.default_rng(520 + weights[0] + 3 * weights[2])
    queries = [rng.integers(97, 100, size=n).astype(np.uint32) for n in (1025, 2048, 2100)]
    queries.append(rng.integers(97, 100, size=300).astype(np.uint32))  # one strip beside them
    choices = [rng.integers(97, 100, size=n).astype(np.uint32) for n in (1025, 1500, 2600)]
    choices += [_gen.mutate(rng, queries[1], 40, alpha_n=3), queries[2].copy(), queries[0][:1024],
                rng.integers(97, 100, size=200).astype(np.uint32), np.zeros(0, np.uint32)]
    check(ctx, oracle, queries, choices, weights, kinds=(0, 3))


def test_per_query_cutoffs(ctx, oracle):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(9)
    weights = (1, 2, 3)
    queries = _gen.random_strings(rng, 40, 0, 64, alpha_n=4) + _gen.random_strings(rng, 3, 70, 200, alpha_n=4)
    choices = _gen.random_strings(rng, 50, 0, 120, alpha_n=4)
    cut = rng.integers(0, 60, size=len(queries)).astype(np.uint64)
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1), weights=weights)
    for kind in (0, 1):
        got = sc.cdist(ctx.strings(choices, width=1), kind, cutoffs_u=cut)
        for i in range(len(queries)):
            exp = oracle.cdist(0, kind, [queries[i]], choices, weights=weights, cutoff_u=int(cut[i]))
            assert np.array_equal(got[i], exp[0]), f"kind {kind} query {i} cutoff {cut[i]}"
    sc.close()


@pytest.mark.parametrize("weights,cutoff", [((1, 2, 3), 0.6), ((2, 2, 3), 0.0), ((3, 1, 2), 0.75)])
def test_topk(ctx, oracle, weights, cutoff):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(11)
    choices = _gen.random_strings(rng, 800, 0, 40, alpha_n=5)
    queries = [_gen.mutate(rng, choices[int(i)], 3, alpha_n=5) for i in rng.integers(0, 800, size=30)]
    queries.append(np.concatenate([choices[0]] * 10)[:150])
    scores = oracle.cdist(0, 3, queries, choices, weights=weights, cutoff_d=cutoff)
    for k in (1, 7):
        exp = expected_topk(scores, cutoff, k)
        sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1), weights=weights)
        gi, gs = sc.topk(ctx.strings(choices, width=1), k, kind=3, cutoff_d=cutoff)
        sc.close()
        assert np.array_equal(gi, exp[0]) and np.array_equal(gs.view(np.uint64), exp[1].view(np.uint64))


def test_limits(ctx):
    import rapidfuzz_amd as rfa

    long_a = np.full(1100, 97, np.uint32)
    sc = rfa.Scorer(ctx, 0, ctx.strings([long_a], width=1), weights=(1, 2, 3))
    assert int(sc.cdist(ctx.strings([long_a], width=1), 0)[0, 0]) == 0  # both longer than 1024 units
    sc.close()
    sc = rfa.Scorer(ctx, 0, ctx.strings([b"abc"]), weights=(1 << 24, 1, 3))
    with pytest.raises(rfa.RfaError):
        sc.cdist(ctx.strings([b"abd"]), 0)
    sc.close()
    with pytest.raises(rfa.RfaError):
        rfa.Server(ctx, ctx.strings([b"abc"]), metric=0, weights=(1, 2, 3))
