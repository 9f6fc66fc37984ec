"""GPU parity of experimental::CachedDamerauLevenshtein (Zhao and Sahni,
DamerauLevenshtein_impl.hpp:36-121) against the oracle, bit for bit: the query
tile kernel for 32- and 64-unit tiles, every kind and cutoff, 8/16/32-bit units,
choices across the 64-unit chunk boundaries, adjacent swaps (the transposition
terms), per-query cutoffs and top-k."""
import numpy as np
import pytest

import _gen
from _topk import expected_topk

pytestmark = pytest.mark.gpu

DL = 18


def _cutoffs(kind):
    if kind in (0, 1):
        return [(None, None), (0, None), (1, None), (2, None), (5, None), (17, None), (200, None)]
    return [(None, None), (None, 0.0), (None, 0.3), (None, 0.5), (None, 0.75), (None, 0.8), (None, 1.0),
            (None, 1.1)]


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def _swapped(rng, s, every=4):
    t = s.copy()
    for p in range(int(rng.integers(0, 3)), len(t) - 1, every):
        t[p], t[p + 1] = t[p + 1], t[p]
    return t


def check(ctx, oracle, queries, choices, width=1, kinds=(0, 1, 2, 3)):
    import rapidfuzz_amd as rfa

    print(f"width {width} nq {len(queries)} nc {len(choices)}", flush=True)
    sc = rfa.Scorer(ctx, DL, ctx.strings(queries, width=width))
    cs = ctx.strings(choices, width=width)
    for kind in kinds:
        for cu, cd in _cutoffs(kind):
            got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd)
            exp = oracle.cdist(DL, kind, queries, choices, cutoff_u=cu, cutoff_d=cd)
            if not _same(got, exp):
                bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
                i, j = bad[0]
                raise AssertionError(f"kind {kind} cutoff {cu}/{cd}: {len(bad)} differ; q{i} (len {len(queries[i])}) "
                                     f"c{j} (len {len(choices[j])}): got {got[i, j]!r} expected {exp[i, j]!r}")
    sc.close()


@pytest.mark.parametrize("alpha", [2, 4, 26])
def test_tile(ctx, oracle, alpha):
    rng = np.random.default_rng(700 + alpha)
    queries = _gen.random_strings(rng, 100, 0, 64, alpha_n=alpha) + [np.zeros(0, np.uint32)]
    queries += [rng.integers(97, 97 + alpha, size=n).astype(np.uint32) for n in (31, 32, 33, 63, 64)]
    choices = _gen.random_strings(rng, 50, 0, 140, alpha_n=alpha) + [np.zeros(0, np.uint32)]
    choices += [_gen.mutate(rng, queries[i], int(rng.integers(0, 4)), alpha_n=alpha) for i in range(20)]
    choices += [_swapped(rng, queries[i]) for i in range(20, 50)]
    choices += [rng.integers(97, 97 + alpha, size=n).astype(np.uint32) for n in (63, 64, 65, 127, 128, 129)]
    check(ctx, oracle, queries, choices)


@pytest.mark.parametrize("width", [2, 4])
def test_tile_wide_units(ctx, oracle, width):
    rng = np.random.default_rng(800 + width)
    lo = 0x4E00 if width == 2 else 0x1F300
    queries = _gen.random_strings(rng, 70, 0, 64, alpha_lo=lo, alpha_n=30)
    choices = _gen.random_strings(rng, 40, 0, 100, alpha_lo=lo, alpha_n=30)
    choices += [_swapped(rng, queries[i], 3) for i in range(25)]
    check(ctx, oracle, queries, choices, width=width)


def test_per_query_cutoffs(ctx, oracle):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(12)
    queries = _gen.random_strings(rng, 45, 0, 64, alpha_n=4)
    choices = _gen.random_strings(rng, 60, 0, 90, alpha_n=4) + [_swapped(rng, q) for q in queries[:20]]
    cut = rng.integers(0, 30, size=len(queries)).astype(np.uint64)
    sc = rfa.Scorer(ctx, DL, ctx.strings(queries, width=1))
    for kind in (0, 1):
        got = sc.cdist(ctx.strings(choices, width=1), kind, cutoffs_u=cut)
        for i in range(len(queries)):
            exp = oracle.cdist(DL, kind, [queries[i]], choices, cutoff_u=int(cut[i]))
            assert np.array_equal(got[i], exp[0]), f"kind {kind} query {i} cutoff {cut[i]}"
    sc.close()


@pytest.mark.parametrize("cutoff", [0.0, 0.7])
def test_topk(ctx, oracle, cutoff):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(13)
    choices = _gen.random_strings(rng, 900, 0, 40, alpha_n=5)
    queries = [_swapped(rng, choices[int(i)], 2) for i in rng.integers(0, 900, size=40)]
    scores = oracle.cdist(DL, 3, queries, choices, cutoff_d=cutoff)
    for k in (1, 6):
        exp = expected_topk(scores, cutoff, k)
        sc = rfa.Scorer(ctx, DL, ctx.strings(queries, width=1))
        gi, gs = sc.topk(ctx.strings(choices, width=1), k, kind=3, cutoff_d=cutoff)
        sc.close()
        assert np.array_equal(gi, exp[0]) and np.array_equal(gs.view(np.uint64), exp[1].view(np.uint64))


@pytest.mark.parametrize("alpha", [3, 26])
def test_long_queries(ctx, oracle, alpha):
    """Queries over 64 units: one lane per pair with the arrays in scratch."""
    rng = np.random.default_rng(900 + alpha)
    queries = [rng.integers(97, 97 + alpha, size=n).astype(np.uint32) for n in (65, 90, 128, 129, 200, 700, 1100)]
    queries += _gen.random_strings(rng, 5, 0, 64, alpha_n=alpha)  # beside the long ones: the tile path
    choices = _gen.random_strings(rng, 40, 0, 300, alpha_n=alpha) + [np.zeros(0, np.uint32)]
    choices += [_swapped(rng, q, 3) for q in queries[:7]]
    choices += [_gen.mutate(rng, q, 6, alpha_n=alpha) for q in queries[:7]]
    check(ctx, oracle, queries, choices)


def test_tiles_with_many_symbols(ctx, oracle):
    """Query tiles of more than 255 distinct 32-bit symbols leave the tile plan."""
    rng = np.random.default_rng(77)
    queries = _gen.random_strings(rng, 64, 20, 64, alpha_lo=0x10000, alpha_n=5000)
    choices = [_swapped(rng, q, 3) for q in queries[:20]] + _gen.random_strings(rng, 30, 0, 80, alpha_lo=0x10000,
                                                                                    alpha_n=5000)
    check(ctx, oracle, queries, choices, width=4, kinds=(0, 3))


def test_limits(ctx):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, DL, ctx.strings([b"abc"]))
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"abd"]), 0, semantics=rfa.MULTI)
    sc.close()
