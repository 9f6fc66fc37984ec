"""GPU parity at the benchmark configurations' own shapes.

* The 4,096 x 65,536 printable-ASCII cdist: queries and choices over the 95
  symbols 0x20-0x7E with lengths 1-64 (tiles of up to 96 PM rows), every scorer
  of that configuration with no cutoff and with normalized_similarity >= 0.8,
  plus integer and double cutoffs that settle pairs by length alone.
* The DNA configuration: ACGT queries of 4,200-10,000 units (the two- and
  four-slot banded instantiations of the block kernel), choices derived by
  1% substitutions / 0.5% insertions / 0.5% deletions and heavier mutations,
  plus unrelated sequences; cutoffs of 5% and 20% of the query length per
  query, and cutoffs d - 1, d, d + 1 around exact distances, through rfa_cdist,
  rfa_cdist_cutoffs and rfa_pairs.
* Every (distance, maximum <= 8192) through the shared normalization epilogue:
  a Hamming cdist of 'a'^i against 'a'^j, i, j = 0..8192, against the closed
  form of tests/_norm.py (pinned against the oracle and the reference by
  tests/test_norm_closed_form.py).
"""
import numpy as np
import pytest

import _gen
import _norm

pytestmark = pytest.mark.gpu

PRINTABLE = dict(alpha_lo=0x20, alpha_n=95)


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def _first_diff(got, exp, queries, choices):
    g = got.view(np.uint64) if got.dtype == np.float64 else got
    e = exp.view(np.uint64) if exp.dtype == np.float64 else exp
    bad = np.argwhere(g != e)
    i, j = bad[0]
    return (f"{len(bad)} mismatches; first q={i} (len {len(queries[i])}) c={j} (len {len(choices[j])}): "
            f"got {got[i, j]!r} expected {exp[i, j]!r}")


@pytest.fixture(scope="module")
def printable_case():
    rng = np.random.default_rng(20261203)
    queries = _gen.random_strings(rng, 512, 1, 64, **PRINTABLE)
    choices = _gen.random_strings(rng, 3072, 1, 64, **PRINTABLE)
    # a quarter of the choices lie near a query, so the 0.8 cutoffs keep some pairs
    for k in range(1024):
        q = queries[int(rng.integers(0, len(queries)))]
        t = _gen.mutate(rng, q, int(rng.integers(0, 13)), **PRINTABLE)
        choices.append(t[:64] if len(t) else np.array([0x41], np.uint32))
    perm = rng.permutation(len(choices))
    return queries, [choices[i] for i in perm]


# (metric, kind, cutoff_u, cutoff_d): the configuration's two runs per scorer first
C2_RUNS = [
    (0, 0, None, None), (0, 3, None, 0.8), (0, 0, 5, None), (0, 2, None, 0.2), (0, 1, 40, None),
    (1, 0, None, None), (1, 3, None, 0.8), (1, 0, 9, None), (1, 2, None, 0.25), (1, 1, 60, None),
    (3, 0, None, None), (3, 3, None, 0.8), (3, 0, 5, None), (3, 2, None, 0.2), (3, 1, 40, None),
    (5, 1, None, None), (5, 3, None, 0.8), (5, 1, None, 0.85), (5, 0, None, 0.1), (5, 2, None, 0.15),
    (4, 1, None, None), (4, 3, None, 0.8),
]


@pytest.mark.parametrize("metric,kind,cu,cd", C2_RUNS)
def test_printable_cdist(ctx, oracle, printable_case, metric, kind, cu, cd):
    import rapidfuzz_amd as rfa

    queries, choices = printable_case
    qs, cs = ctx.strings(queries, width=1), ctx.strings(choices, width=1)
    sc = rfa.Scorer(ctx, metric, qs, prefix_weight=0.1)
    got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd)
    exp = oracle.cdist_mt(metric, kind, queries, choices, cutoff_u=cu, cutoff_d=cd)
    sc.close()
    assert _same(got, exp), f"metric {metric} kind {kind} cutoff {cu}/{cd}: " + _first_diff(got, exp, queries,
                                                                                            choices)


@pytest.mark.parametrize("metric", [0, 1, 3])
def test_printable_per_query_cutoffs(ctx, oracle, printable_case, metric):
    """Per-query integer cutoffs on the printable tiles (one cutoff per lane of a tile)."""
    import rapidfuzz_amd as rfa

    queries, choices = printable_case
    queries, choices = queries[:192], choices[:1024]
    rng = np.random.default_rng(metric)
    cuts = rng.integers(0, 30, size=len(queries)).astype(np.uint64)
    sc = rfa.Scorer(ctx, metric, ctx.strings(queries, width=1))
    got = sc.cdist(ctx.strings(choices, width=1), 0, cutoffs_u=cuts)
    exp = oracle.cdist_mt(metric, 0, queries, choices, cutoffs_u=cuts)
    assert _same(got, exp), _first_diff(got, exp, queries, choices)


ACGT = np.array([65, 67, 71, 84], dtype=np.uint32)


def _dna_mutate(rng, s, sub, ins, dele):
    """Per-position substitution / insertion / deletion with the given rates (ACGT)."""
    code = {int(v): k for k, v in enumerate(ACGT)}
    out = []
    for x in s:
        r = rng.random()
        if r < dele:
            continue
        if r < dele + sub:
            out.append(int(ACGT[(code[int(x)] + int(rng.integers(1, 4))) % 4]))
        else:
            out.append(int(x))
        if rng.random() < ins:
            out.append(int(ACGT[int(rng.integers(0, 4))]))
    return np.array(out, dtype=np.uint32)


def _acgt(rng, n):
    return ACGT[rng.integers(0, 4, size=n)]


def _dna_case(seed, lengths):
    rng = np.random.default_rng(seed)
    queries = [_acgt(rng, L) for L in lengths]
    choices = []
    for q in queries:
        choices.append(_dna_mutate(rng, q, 0.01, 0.005, 0.005))
        choices.append(_dna_mutate(rng, q, 0.01, 0.005, 0.005))
        choices.append(_dna_mutate(rng, q, 0.03, 0.01, 0.01))
        choices.append(_dna_mutate(rng, q, 0.09, 0.03, 0.03))
        choices.append(_dna_mutate(rng, q[: len(q) * 9 // 10], 0.01, 0.005, 0.005))
    for L in (1000, 4600, 9900):
        choices.append(_acgt(rng, L))
    return queries, choices


@pytest.fixture(scope="module")
def dna_cases(oracle):
    cases = {}
    for name, seed, lengths in (("four_slot", 31, (4500, 7300, 10000)), ("two_slot", 32, (4200, 6100, 8000))):
        queries, choices = _dna_case(seed, lengths)
        full = oracle.cdist_mt(0, 0, queries, choices, threads=len(queries))
        cases[name] = (queries, choices, full)
    return cases


@pytest.mark.parametrize("name", ["four_slot", "two_slot"])
def test_dna_per_query_cutoffs(ctx, oracle, dna_cases, name):
    """Cutoffs floor(0.05 m) and floor(0.2 m) per query (the q-gram filter is active at 5%, where
    7k + 7 < m, and off at 20%), against the oracle run with each query's cutoff."""
    import rapidfuzz_amd as rfa

    queries, choices, full = dna_cases[name]
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1))
    cs = ctx.strings(choices, width=1)
    for frac in (0.05, 0.20):
        cuts = np.array([int(frac * len(q)) for q in queries], dtype=np.uint64)
        got = sc.cdist(cs, 0, cutoffs_u=cuts)
        exp = oracle.cdist_mt(0, 0, queries, choices, cutoffs_u=cuts)
        assert _same(got, exp), f"{name} {frac}: " + _first_diff(got, exp, queries, choices)
        # the oracle's capped values equal the exact distances cut at each query's cutoff
        assert np.array_equal(exp, np.where(full <= cuts[:, None], full, cuts[:, None] + 1))
        assert np.count_nonzero(full <= cuts[:, None]) > 0


@pytest.mark.parametrize("name", ["four_slot", "two_slot"])
def test_dna_cutoffs_at_distances(ctx, dna_cases, name):
    """One cutoff for the whole launch, at d - 1, d and d + 1 of exact distances."""
    import rapidfuzz_amd as rfa

    queries, choices, full = dna_cases[name]
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1))
    cs = ctx.strings(choices, width=1)
    ds = sorted({int(d) for d in full.ravel() if d < 3000})
    cutoffs = sorted({max(0, d + e) for d in ds[::2] for e in (-1, 0, 1)} | {0, 50})
    for cu in cutoffs:
        got = sc.cdist(cs, 0, cutoff_u=cu)
        exp = np.where(full <= cu, full, cu + 1).astype(np.uint64)
        assert np.array_equal(got, exp), f"{name} cutoff {cu}: {np.count_nonzero(got != exp)} differ"


def test_dna_pairs(ctx, oracle, dna_cases):
    """The same pairs through rfa_pairs (the bucketed schedule; the plain one-pair-per-lane schedule
    takes patterns of at most 4,096 units), every query against every choice."""
    import rapidfuzz_amd as rfa

    queries, choices, full = dna_cases["four_slot"]
    a = [queries[i] for i in range(len(queries)) for _ in choices]
    b = [c for _ in queries for c in choices]
    metrics = np.zeros(len(a), np.int32)
    A, B = ctx.strings(a, width=1), ctx.strings(b, width=1)
    flat = full.ravel()
    exp500 = oracle.pairs(metrics, a, b, kind=0, cutoff_u=500)
    assert np.array_equal(exp500, np.where(flat <= 500, flat, 501))
    for cu in (0, 225, 500, 1460, 2000, int(np.median(flat)), rfa.INT64_MAX):
        exp = np.where(flat <= cu, flat, cu + 1).astype(np.uint64) if cu != rfa.INT64_MAX else flat
        got = rfa.pairs(ctx, A, B, metrics, kind=0, cutoff_u=cu, schedule=rfa.SCHEDULE_BUCKETED)
        assert np.array_equal(got, exp), f"cutoff {cu}: {np.count_nonzero(got != exp)} differ"


NMAX = 8192


@pytest.fixture(scope="module")
def hamming_grid(ctx):
    import rapidfuzz_amd as rfa

    strs = [np.full(i, 97, dtype=np.uint8) for i in range(NMAX + 1)]
    s = ctx.strings(strs, width=1)
    sc = rfa.Scorer(ctx, 7, s)
    yield sc, s
    sc.close()
    s.close()


def _grid_rows(r0, r1):
    i = np.arange(r0, r1, dtype=np.int64)[:, None]
    j = np.arange(NMAX + 1, dtype=np.int64)[None, :]
    return np.abs(i - j).astype(np.uint64), np.maximum(i, j).astype(np.uint64)


@pytest.mark.parametrize("kind,c", [(2, 1.0), (2, 0.2), (2, 1.0 / 3.0), (2, 0.05),
                                    (3, 0.0), (3, 0.8), (3, 2.0 / 3.0), (3, 0.95)])
def test_normalization_exhaustive(hamming_grid, kind, c):
    """Every (d, maximum) with maximum <= 8192 through the normalized epilogue."""
    sc, s = hamming_grid
    got = sc.cdist(s, kind, cutoff_d=c)
    f = _norm.norm_distance if kind == 2 else _norm.norm_similarity
    step = 1024
    for r0 in range(0, NMAX + 1, step):
        r1 = min(NMAX + 1, r0 + step)
        d, M = _grid_rows(r0, r1)
        exp = f(d, M, c)
        g = got[r0:r1]
        if not np.array_equal(g.view(np.uint64), exp.view(np.uint64)):
            bad = np.argwhere(g.view(np.uint64) != exp.view(np.uint64))
            i, j = bad[0]
            raise AssertionError(f"kind {kind} cutoff {c}: {len(bad)} differ in rows {r0}-{r1}; first i={r0 + i} "
                                 f"j={j}: got {g[i, j]!r} expected {exp[i, j]!r}")


@pytest.mark.parametrize("metric", [0, 1, 3])
@pytest.mark.parametrize("kind,c", [(2, 0.2), (3, 0.8), (3, 2.0 / 3.0)])
def test_normalization_tile_grid(ctx, oracle, metric, kind, c):
    """The query-tile kernel's epilogue over 'a'^i x 'a'^j, i, j <= 64 (both word widths)."""
    import rapidfuzz_amd as rfa

    strs = [np.full(i, 97, dtype=np.uint32) for i in range(65)]
    sc = rfa.Scorer(ctx, metric, ctx.strings(strs, width=1))
    got = sc.cdist(ctx.strings(strs, width=1), kind, cutoff_d=c)
    exp = oracle.cdist(metric, kind, strs, strs, cutoff_d=c)
    assert _same(got, exp)


@pytest.mark.parametrize("metric", [0, 1, 3, 6])
@pytest.mark.parametrize("kind,c", [(2, 0.2), (2, 1.0 / 3.0), (2, 1.0), (3, 0.8), (3, 2.0 / 3.0), (3, 0.0),
                                    (3, 1.0)])
def test_normalization_levenshtein_family(ctx, metric, kind, c):
    """The normalized epilogue of Levenshtein / OSA (d = |i - j|, maximum max(i, j)), Indel
    (maximum i + j) and ratio over 'a'^i x 'a'^j, i, j <= 1024: every (d, maximum) those give,
    through the tile kernels (<= 64 units) and the block kernel (longer queries)."""
    import rapidfuzz_amd as rfa

    if metric == 6 and kind != 3:
        pytest.skip("ratio is a similarity")
    n = 1024
    strs = [np.full(i, 97, dtype=np.uint8) for i in range(n + 1)]
    s = ctx.strings(strs, width=1)
    sc = rfa.Scorer(ctx, metric, s)
    i = np.arange(n + 1, dtype=np.int64)
    d = np.abs(i[:, None] - i[None, :]).astype(np.uint64)
    M = (i[:, None] + i[None, :] if metric in (1, 6) else np.maximum(i[:, None], i[None, :])).astype(np.uint64)
    if metric == 6:
        got = sc.cdist(s, 1, cutoff_d=c * 100)
        exp = _norm.norm_similarity(d, M, (c * 100) / 100) * 100
    else:
        got = sc.cdist(s, kind, cutoff_d=c)
        exp = (_norm.norm_distance if kind == 2 else _norm.norm_similarity)(d, M, c)
    sc.close()
    s.close()
    if not np.array_equal(got.view(np.uint64), exp.view(np.uint64)):
        bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
        a, b = bad[0]
        raise AssertionError(f"{len(bad)} differ; first i={a} j={b}: got {got[a, b]!r} expected {exp[a, b]!r}")


@pytest.mark.parametrize("metric", [1, 2, 6])
def test_lcs_cutoff_sweep(ctx, oracle, printable_case, metric):
    """Indel, LCSseq and ratio at many cutoffs of every kind on the printable tiles: the LCS
    recurrence stops once no lane can reach the LCS its cutoff needs, and the results must not
    change.  Choices near a query put exact values on both sides of each threshold."""
    import rapidfuzz_amd as rfa

    queries, choices = printable_case
    queries, choices = queries[:128], choices[:1536]
    qs, cs = ctx.strings(queries, width=1), ctx.strings(choices, width=1)
    sc = rfa.Scorer(ctx, metric, qs)
    runs = []
    if metric == 6:
        runs = [(1, None, c) for c in (0.0, 30.0, 50.0, 70.0, 85.0, 95.0, 100.0)]
    else:
        runs += [(0, c, None) for c in (0, 1, 3, 8, 15, 30, 60)]
        runs += [(1, c, None) for c in (1, 5, 10, 20, 40)]
        runs += [(2, None, c) for c in (0.05, 0.15, 0.3, 0.6)]
        runs += [(3, None, c) for c in (0.4, 0.7, 0.85, 0.95)]
    for kind, cu, cd in runs:
        got = sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd)
        exp = oracle.cdist_mt(metric, kind, queries, choices, cutoff_u=cu, cutoff_d=cd)
        assert _same(got, exp), f"kind {kind} cutoff {cu}/{cd}: " + _first_diff(got, exp, queries, choices)
    if metric != 6:
        rng = np.random.default_rng(metric)
        for kind in (0, 1):
            cuts = rng.integers(0, 40, size=len(queries)).astype(np.uint64)
            got = sc.cdist(cs, kind, cutoffs_u=cuts)
            exp = oracle.cdist_mt(metric, kind, queries, choices, cutoffs_u=cuts)
            assert _same(got, exp), f"per-query kind {kind}: " + _first_diff(got, exp, queries, choices)
    sc.close()
