"""Randomized parity sweeps (GPU against the oracle, bit for bit) over the weighted
Levenshtein, Damerau-Levenshtein and editops kernels: many weight triples, query
lengths on both sides of the 32- and 64-row tiles, long queries beside short ones, and
a large mixed editops batch.  Seeds are fixed; each case prints what it runs."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


@pytest.mark.parametrize("seed", range(6))
def test_weighted_levenshtein_random_weights(ctx, oracle, seed):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(1000 + seed)
    queries = _gen.random_strings(rng, 40, 0, 64, alpha_n=3) + _gen.random_strings(rng, 3, 65, 200, alpha_n=3)
    choices = _gen.random_strings(rng, 30, 0, 120, alpha_n=3)
    choices += [_gen.mutate(rng, q, int(rng.integers(0, 6)), alpha_n=3) for q in queries[:15]]
    cs = ctx.strings(choices, width=1)
    for _ in range(4):
        w = tuple(int(x) for x in rng.integers(0, 9, size=3))
        kind = int(rng.integers(0, 4))
        cu = int(rng.integers(0, 40)) if kind < 2 else None
        cd = float(rng.choice([0.0, 0.5, 0.8])) if kind >= 2 else None
        print(f"seed {seed} weights {w} kind {kind} cutoff {cu}/{cd}", flush=True)
        sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1), weights=w)
        got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd)
        exp = oracle.cdist(0, kind, queries, choices, weights=w, cutoff_u=cu, cutoff_d=cd)
        sc.close()
        assert _same(got, exp), f"weights {w} kind {kind}"


@pytest.mark.parametrize("seed", range(4))
def test_damerau_tile_boundaries(ctx, oracle, seed):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(2000 + seed)
    alpha = int(rng.integers(2, 6))
    lens = [0, 1, 2, 3, 4, 5, 31, 32, 33, 47, 48, 49, 63, 64, 65, 66, 100]
    queries = [rng.integers(97, 97 + alpha, size=n).astype(np.uint32) for n in lens for _ in range(3)]
    choices = [rng.integers(97, 97 + alpha, size=int(rng.integers(0, 140))).astype(np.uint32) for _ in range(40)]
    choices += [q[::-1].copy() for q in queries[::5]]
    print(f"seed {seed} alpha {alpha}", flush=True)
    sc = rfa.Scorer(ctx, 18, ctx.strings(queries, width=1))
    for kind, cu, cd in ((0, None, None), (1, 3, None), (3, None, 0.6)):
        got = sc.cdist(ctx.strings(choices, width=1), kind, cutoff_u=cu, cutoff_d=cd)
        exp = oracle.cdist(18, kind, queries, choices, cutoff_u=cu, cutoff_d=cd)
        assert _same(got, exp), f"kind {kind}"
    sc.close()


@pytest.mark.parametrize("metric", [0, 2, 7])
def test_editops_large_batch(ctx, oracle, metric):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(3000 + metric)
    a, b = [], []
    for _ in range(3000):
        s1 = rng.integers(97, 101, size=int(rng.integers(0, 60))).astype(np.uint32)
        pre = rng.integers(97, 101, size=int(rng.integers(0, 20))).astype(np.uint32)
        a.append(np.concatenate([pre, s1]))
        b.append(np.concatenate([pre, _gen.mutate(rng, s1, int(rng.integers(0, 12)))]))
    print(f"metric {metric} pairs {len(a)}", flush=True)
    off, ops = rfa.editops(ctx, ctx.strings(a, width=1), ctx.strings(b, width=1), metric)
    for i in range(len(a)):
        got = ops[int(off[i]):int(off[i + 1])]
        exp = oracle.editops(metric, a[i], b[i])
        assert np.array_equal(got["type"], exp[0]) and np.array_equal(got["src_pos"], exp[1]) and \
            np.array_equal(got["dest_pos"], exp[2]), f"pair {i}"
