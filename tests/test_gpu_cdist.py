"""GPU parity: rfa_cdist against the oracle, bit for bit (integers and doubles)."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu

METRICS = [0, 1, 2, 3, 4, 5, 6]  # Levenshtein, Indel, LCSseq, OSA, Jaro, JaroWinkler, ratio
KINDS = [0, 1, 2, 3]


def _kinds(metric):
    return [1] if metric == 6 else KINDS


def _cutoffs(metric, kind):
    """(cutoff_u, cutoff_d) pairs exercising the reference's cutoff branches."""
    int_kind = metric in (0, 1, 2, 3) and kind in (0, 1)
    if int_kind:
        return [(None, None), (0, None), (1, None), (3, None), (7, None), (20, None)]
    if metric == 6:
        return [(None, 0.0), (None, 50.0), (None, 70.0), (None, 90.0), (None, 100.0)]
    return [(None, None), (None, 0.0), (None, 0.2), (None, 0.7), (None, 0.8), (None, 0.85), (None, 1.0),
            (None, 1.1)]


def same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def check(ctx, oracle, metric, queries, choices, width=None, weights=(1, 1, 1), semantics=0, kinds=None):
    import rapidfuzz_amd as rfa

    qs = ctx.strings(queries, width=width)
    cs = ctx.strings(choices, width=width)
    sc = rfa.Scorer(ctx, metric, qs, weights=weights, prefix_weight=0.1)
    for kind in kinds or _kinds(metric):
        for cu, cd in _cutoffs(metric, kind):
            got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd, semantics=semantics)
            exp = oracle.cdist(metric, kind, queries, choices, weights=weights, cutoff_u=cu, cutoff_d=cd,
                               multi=bool(semantics))
            if not same(got, exp):
                bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64)) if got.dtype == np.float64 \
                    else np.argwhere(got != exp)
                i, j = bad[0]
                raise AssertionError(
                    f"metric {metric} kind {kind} sem {semantics} cutoff {cu}/{cd}: {len(bad)} mismatches; "
                    f"first q={i} (len {len(queries[i])}) c={j} (len {len(choices[j])}): "
                    f"got {got[i, j]!r} expected {exp[i, j]!r}")


@pytest.mark.parametrize("metric", METRICS)
def test_short_random(ctx, oracle, metric):
    rng = np.random.default_rng(100 + metric)
    queries = _gen.random_strings(rng, 70, 0, 64, alpha_n=5)
    choices = _gen.random_strings(rng, 40, 0, 70, alpha_n=5)
    choices += [_gen.mutate(rng, queries[i], int(rng.integers(0, 6)), alpha_n=5) for i in range(20)]
    check(ctx, oracle, metric, queries, choices)


@pytest.mark.parametrize("metric", METRICS)
def test_edges(ctx, oracle, metric):
    e = _gen.edge_strings()
    check(ctx, oracle, metric, e, e)


@pytest.mark.parametrize("metric", METRICS)
def test_long_patterns(ctx, oracle, metric):
    rng = np.random.default_rng(200 + metric)
    queries = _gen.random_strings(rng, 6, 65, 300, alpha_n=4)
    choices = _gen.random_strings(rng, 8, 0, 300, alpha_n=4)
    choices += [_gen.mutate(rng, queries[i], int(rng.integers(0, 30)), alpha_n=4) for i in range(6)]
    check(ctx, oracle, metric, queries, choices)


@pytest.mark.parametrize("metric", METRICS)
def test_multi_semantics(ctx, oracle, metric):
    rng = np.random.default_rng(300 + metric)
    queries = _gen.random_strings(rng, 20, 0, 64, alpha_n=3)
    choices = _gen.random_strings(rng, 20, 0, 80, alpha_n=3)
    check(ctx, oracle, metric, queries, choices, semantics=1)


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("metric", [0, 1, 3, 5])
def test_wide_units(ctx, oracle, metric, width):
    rng = np.random.default_rng(400 + metric + width)
    hi = 60000 if width == 2 else 3_000_000_000
    pool = rng.integers(0, hi, size=40).astype(np.uint32)
    def pick(n, lo, up):
        return [pool[rng.integers(0, len(pool), size=int(rng.integers(lo, up + 1)))] for _ in range(n)]
    queries = pick(30, 0, 64) + pick(3, 65, 200)
    choices = pick(20, 0, 150)
    check(ctx, oracle, metric, queries, choices, width=width)


@pytest.mark.parametrize("weights", [(2, 2, 2), (1, 1, 2), (3, 3, 7), (0, 0, 5)])
def test_levenshtein_weights(ctx, oracle, weights):
    rng = np.random.default_rng(7)
    queries = _gen.random_strings(rng, 20, 0, 80, alpha_n=3)
    choices = _gen.random_strings(rng, 20, 0, 80, alpha_n=3)
    check(ctx, oracle, 0, queries, choices, weights=weights)


def test_unsupported_weights(ctx):
    import rapidfuzz_amd as rfa

    qs = ctx.strings([b"abc"])
    with pytest.raises(rfa.RfaError):
        rfa.Scorer(ctx, 0, qs, weights=(1, 2, 3))


def test_ratio_rejects_distance(ctx):
    import rapidfuzz_amd as rfa

    qs = ctx.strings([b"abc"])
    cs = ctx.strings([b"abd"])
    sc = rfa.Scorer(ctx, rfa.RATIO, qs)
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(cs, rfa.DISTANCE)
