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
    int_kind = metric in (0, 1, 2, 3, 7, 8, 9) and kind in (0, 1)
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


def test_general_weights_accepted(ctx, oracle):
    """Weights off the uniform and Indel-like forms take the weighted Wagner-Fischer kernels."""
    check(ctx, oracle, 0, [b"abc", b"", b"kitten"], [b"abd", b"sitting", b""], width=1, weights=(1, 2, 3))


def test_ratio_rejects_distance(ctx):
    import rapidfuzz_amd as rfa

    qs = ctx.strings([b"abc"])
    cs = ctx.strings([b"abd"])
    sc = rfa.Scorer(ctx, rfa.RATIO, qs)
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(cs, rfa.DISTANCE)


@pytest.mark.parametrize("nq", [1, 5, 16])
@pytest.mark.parametrize("metric,weights", [(0, (1, 1, 1)), (0, (2, 2, 2)), (0, (1, 1, 2)), (1, (1, 1, 1)),
                                            (2, (1, 1, 1)), (6, (1, 1, 1))])
def test_few_queries_one_pair_per_lane(ctx, oracle, nq, metric, weights):
    """At most 16 one-byte queries of <= 256 units: one pair per lane over the length-sorted set."""
    rng = np.random.default_rng(40 + nq + 3 * metric + weights[2])
    queries = _gen.random_strings(rng, nq - 1, 0, 256, alpha_n=6) + [np.zeros(0, np.uint32)]
    choices = _gen.random_strings(rng, 150, 0, 300, alpha_n=6)
    choices += [_gen.mutate(rng, q, int(rng.integers(0, 6)), alpha_n=6) for q in queries]
    # MultiLevenshtein takes weights {1, 1, <= 2} only
    multi_ok = metric != 0 or (weights[0] == 1 and weights[1] == 1 and weights[2] <= 2)
    for semantics in ((0, 1) if multi_ok else (0,)):
        check(ctx, oracle, metric, queries, choices, width=1, weights=weights, semantics=semantics)


@pytest.mark.parametrize("weights", [(1, 1, 0), (1, 1, 1), (1, 1, 2)])
def test_multi_levenshtein_weights(ctx, oracle, weights):
    """MultiLevenshtein: the uniform distance whatever the weights, maximum() from the weights;
    both the tile path (> 16 queries) and the one-pair-per-lane path (<= 16 queries)."""
    rng = np.random.default_rng(11 + weights[2])
    for nq in (40, 7):
        queries = _gen.random_strings(rng, nq, 0, 64, alpha_n=4)
        choices = _gen.random_strings(rng, 40, 0, 90, alpha_n=4)
        check(ctx, oracle, 0, queries, choices, width=1, weights=weights, semantics=1)


def test_multi_levenshtein_rejects_weights(ctx):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, 0, ctx.strings([b"abc"]), weights=(2, 2, 2))
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"abd"]), rfa.DISTANCE, semantics=rfa.MULTI)
    sc0 = rfa.Scorer(ctx, 0, ctx.strings([b"abc"]), weights=(1, 1, 0))  # Multi-only weights
    # Cached scoring with them runs the general-weight recurrence: a free replacement
    assert sc0.cdist(ctx.strings([b"abd"]), rfa.DISTANCE)[0, 0] == 0


def _dna_case(seed, nq=6, nc=40, lo=300, hi=1500):
    rng = np.random.default_rng(seed)
    queries = [rng.integers(65, 69, size=int(rng.integers(lo, hi))).astype(np.uint32) for _ in range(nq)]
    choices = []
    for i in range(nc):
        q = queries[i % nq]
        kind = i % 4
        if kind == 0:
            choices.append(_gen.mutate(rng, q, int(rng.integers(0, len(q) // 20 + 1)), alpha_lo=65, alpha_n=4))
        elif kind == 1:
            choices.append(_gen.mutate(rng, q, int(rng.integers(len(q) // 20, len(q) // 5 + 2)), alpha_lo=65,
                                       alpha_n=4))
        elif kind == 2:
            choices.append(rng.integers(65, 69, size=len(q)).astype(np.uint32))
        else:
            choices.append(rng.integers(65, 69, size=max(1, len(q) + int(rng.integers(-40, 40)))).astype(np.uint32))
    return queries, choices


def test_banded_levenshtein_cutoffs(ctx, oracle):
    """Levenshtein distance with cutoffs on long patterns: the banded block recurrence must equal the
    oracle at every cutoff, in particular where the cutoff is at or next to the distance."""
    import rapidfuzz_amd as rfa

    queries, choices = _dna_case(3)
    full = oracle.cdist(0, 0, queries, choices)
    cutoffs = sorted({0, 1, 5, 20, 60, 150, 400, 900} | {int(d) + e for d in full.ravel()[::7] for e in (-1, 0, 1)
                                                         if int(d) + e >= 0})
    qs, cs = ctx.strings(queries, width=1), ctx.strings(choices, width=1)
    sc = rfa.Scorer(ctx, 0, qs)
    for cu in cutoffs:
        for semantics in (0, 1):
            got = sc.cdist(cs, 0, cutoff_u=cu, semantics=semantics)
            exp = np.where(full <= cu, full, cu + 1).astype(np.uint64)
            assert np.array_equal(got, exp), f"cutoff {cu} semantics {semantics}: {np.count_nonzero(got != exp)} differ"


@pytest.mark.parametrize("width", [1, 2, 4])
def test_banded_qgram_filter_bound(ctx, oracle, width):
    """The q-gram count filter in front of the banded recurrence must never reject a pair whose
    distance is within the cutoff.  Substitutions every 8th unit destroy the most 8-grams per edit,
    so distances d with cutoffs d - 1, d, d + 1 sit exactly on the filter's bound; unrelated strings
    and random mutations cover the rejecting side.  16/32-bit units take the hashed q-gram path."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(90 + width)
    base = {1: 65, 2: 40000, 4: 3000000}[width]
    queries, choices = [], []
    for L in (120, 333, 800, 1500):
        q = (rng.integers(0, 4, size=L) + base).astype(np.uint32)
        queries.append(q)
        for step in (8, 9, 16, 40):
            t = q.copy()
            t[::step] = base + (t[::step] - base + 1) % 4
            choices.append(t)
        choices.append(_gen.mutate(rng, q - base, int(rng.integers(1, L // 30 + 2)), alpha_lo=0, alpha_n=4)
                       .astype(np.uint32) + base)
        choices.append((rng.integers(0, 4, size=L + int(rng.integers(-5, 6))) + base).astype(np.uint32))
    full = oracle.cdist(0, 0, queries, choices)
    cutoffs = sorted({0, 2, 7, 15, 40} | {int(d) + e for d in full.ravel() for e in (-1, 0, 1) if 0 <= int(d) + e < 400})
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=width))
    cs = ctx.strings(choices, width=width)
    for cu in cutoffs:
        got = sc.cdist(cs, 0, cutoff_u=cu)
        exp = np.where(full <= cu, full, cu + 1).astype(np.uint64)
        assert np.array_equal(got, exp), f"cutoff {cu}: {np.count_nonzero(got != exp)} differ"


def test_banded_levenshtein_pairs(ctx, oracle):
    """The same through rfa_pairs (the mixed-metric block kernel), both schedules."""
    import rapidfuzz_amd as rfa

    queries, choices = _dna_case(4, nq=10, nc=60, lo=100, hi=2500)
    a = [queries[i % len(queries)] for i in range(len(choices))]
    b = choices
    metrics = np.zeros(len(a), np.int32)
    A, B = ctx.strings(a, width=1), ctx.strings(b, width=1)
    for cu in (0, 3, 30, 120, 500):
        exp = oracle.pairs(metrics, a, b, kind=0, cutoff_u=cu)
        for schedule in (0, 1):
            got = rfa.pairs(ctx, A, B, metrics, kind=0, cutoff_u=cu, schedule=schedule)
            assert np.array_equal(got, exp), f"cutoff {cu} schedule {schedule}"


@pytest.mark.parametrize("metric", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", ["tile", "few", "block"])
def test_per_query_cutoffs(ctx, oracle, metric, shape):
    """rfa_cdist_cutoffs: query q scored with its own cutoff, as a loop of Cached<Metric>(q) calls."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(70 + metric + 5 * len(shape))
    if shape == "tile":
        queries = _gen.random_strings(rng, 40, 0, 64, alpha_n=4)
    elif shape == "few":
        queries = _gen.random_strings(rng, 6, 0, 200, alpha_n=4)
    else:
        queries = _gen.random_strings(rng, 12, 100, 700, alpha_n=4)
    choices = _gen.random_strings(rng, 30, 0, 300, alpha_n=4)
    choices += [_gen.mutate(rng, q, int(rng.integers(0, 10))) for q in queries[:10]]
    sc = rfa.Scorer(ctx, metric, ctx.strings(queries, width=1))
    cs = ctx.strings(choices, width=1)
    for kind in (0, 1):
        for semantics in (0, 1):
            cuts = rng.integers(0, 60, size=len(queries)).astype(np.uint64)
            got = sc.cdist(cs, kind, semantics=semantics, cutoffs_u=cuts)
            exp = np.stack([oracle.cdist(metric, kind, [queries[i]], choices, cutoff_u=int(cuts[i]),
                                         multi=bool(semantics))[0] for i in range(len(queries))])
            assert np.array_equal(got, exp), f"kind {kind} semantics {semantics}: {np.count_nonzero(got != exp)} differ"


@pytest.mark.parametrize("metric", [7, 8, 9])
@pytest.mark.parametrize("width", [1, 2, 4])
def test_positional_metrics(ctx, oracle, metric, width):
    """Hamming (padded), Prefix and Postfix: every kind and cutoff, queries and choices of 0-150
    units sharing prefixes and suffixes, 8/16/32-bit units, per-query cutoffs."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(400 + metric * 3 + width)
    base = {1: 97, 2: 50000, 4: 2000000}[width]
    stems = [rng.integers(0, 3, size=int(rng.integers(0, 150))).astype(np.uint32) + base for _ in range(8)]

    def variant(s):
        t = s.copy()
        if len(t) and rng.random() < 0.7:
            t[rng.integers(0, len(t), size=int(rng.integers(1, 4)))] = base + 3
        if rng.random() < 0.5:
            t = np.concatenate([t, rng.integers(0, 3, size=int(rng.integers(0, 9))).astype(np.uint32) + base])
        if rng.random() < 0.3:
            t = np.concatenate([rng.integers(0, 3, size=int(rng.integers(1, 5))).astype(np.uint32) + base, t])
        return t

    queries = [variant(stems[i % 8]) for i in range(70)] + [np.zeros(0, np.uint32)]
    choices = [variant(stems[i % 8]) for i in range(150)] + [np.zeros(0, np.uint32), stems[0][:1]]
    sc = rfa.Scorer(ctx, metric, ctx.strings(queries, width=width))
    cs = ctx.strings(choices, width=width)
    for kind in (0, 1, 2, 3):
        cuts = [(None, None)]
        if kind in (0, 1):
            cuts += [(c, None) for c in (0, 1, 3, 10, 60)]
        else:
            cuts += [(None, c) for c in (0.0, 0.2, 0.5, 0.8, 1.0)]
        for cu, cd in cuts:
            kw = {}
            if cu is not None:
                kw["cutoff_u"] = cu
            if cd is not None:
                kw["cutoff_d"] = cd
            got = sc.cdist(cs, kind, **kw)
            exp = oracle.cdist(metric, kind, queries, choices, **kw)
            if got.dtype == np.float64:
                same = np.array_equal(got.view(np.uint64), exp.view(np.uint64))
            else:
                same = np.array_equal(got, exp)
            assert same, f"kind {kind} cutoff {cu}/{cd}: {np.count_nonzero(got != exp)} differ"
    cuts = rng.integers(0, 40, size=len(queries)).astype(np.uint64)
    for kind in (0, 1):
        got = sc.cdist(cs, kind, cutoffs_u=cuts)
        exp = np.stack([oracle.cdist(metric, kind, [queries[i]], choices, cutoff_u=int(cuts[i]))[0]
                        for i in range(len(queries))])
        assert np.array_equal(got, exp), f"per-query cutoffs kind {kind}"
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(cs, 0, semantics=1)
