"""GPU parity of the diagonal band kernel (cdist of long 8-bit queries under a
Levenshtein distance cutoff): every result against the oracle, and the band kernel
against the one-wave-per-pair block kernel (RFA_OPTION_CDIST_BAND = 0)."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu


def _case(seed, nq, nc, lo, hi, alpha_n):
    """Queries of lo..hi units; choices that are mutated queries (light to heavy edits),
    truncated or extended queries (length differences near the cutoffs) and random strings
    of about a query's length."""
    rng = np.random.default_rng(seed)
    queries = _gen.random_strings(rng, nq, lo, hi, alpha_lo=65, alpha_n=alpha_n)
    choices = []
    for i in range(nc):
        q = queries[int(rng.integers(0, nq))]
        kind = i % 5
        if kind == 0:
            choices.append(_gen.mutate(rng, q, int(rng.integers(0, len(q) // 15 + 1)), alpha_lo=65, alpha_n=alpha_n))
        elif kind == 1:
            choices.append(_gen.mutate(rng, q, int(rng.integers(len(q) // 15, len(q) // 3 + 2)), alpha_lo=65,
                                       alpha_n=alpha_n))
        elif kind == 2:
            cut = int(rng.integers(0, len(q) // 4 + 1))
            choices.append(q[cut:] if rng.integers(0, 2) else q[: len(q) - cut])
        elif kind == 3:
            extra = rng.integers(65, 65 + alpha_n, size=int(rng.integers(1, len(q) // 4 + 2))).astype(np.uint32)
            choices.append(np.concatenate([extra, q]) if rng.integers(0, 2) else np.concatenate([q, extra]))
        else:
            n = max(1, len(q) + int(rng.integers(-len(q) // 5, len(q) // 5 + 1)))
            choices.append(rng.integers(65, 65 + alpha_n, size=n).astype(np.uint32))
    choices.append(np.zeros(0, dtype=np.uint32))
    return queries, choices


def _run(ctx, queries, choices, cuts, band):
    import rapidfuzz_amd as rfa

    ctx.set_option(rfa.OPTION_CDIST_BAND, 1 if band else 0)
    try:
        sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1))
        return sc.cdist(ctx.strings(choices, width=1), 0, cutoffs_u=cuts)
    finally:
        ctx.set_option(rfa.OPTION_CDIST_BAND, 1)


def _check(ctx, oracle, queries, choices, cuts, full):
    exp = np.where(full <= cuts[:, None], full, cuts[:, None] + 1).astype(np.uint64)
    got = _run(ctx, queries, choices, cuts, True)
    bad = np.argwhere(got != exp)
    if len(bad):
        i, j = bad[0]
        raise AssertionError(f"{len(bad)} differ; first q={i} (len {len(queries[i])}, k {cuts[i]}) c={j} "
                             f"(len {len(choices[j])}): got {got[i, j]} expected {exp[i, j]} (distance {full[i, j]})")
    assert np.array_equal(_run(ctx, queries, choices, cuts, False), exp)


@pytest.mark.parametrize("alpha_n", [2, 4, 26])
def test_band_random_cutoffs(ctx, oracle, alpha_n):
    """One-word and multi-word windows (k from 0 to about m / 2), every length difference up to the
    cutoff, empty choices; per-query cutoffs at, next to and away from the distances."""
    queries, choices = _case(400 + alpha_n, 10, 160, 65, 1200, alpha_n)
    full = oracle.cdist_mt(0, 0, queries, choices)
    rng = np.random.default_rng(alpha_n)
    for rep in range(4):
        cuts = []
        for i, q in enumerate(queries):
            row = full[i]
            if rep == 0:
                k = int(rng.integers(0, 64))
            elif rep == 1:
                k = int(rng.integers(0, len(q) // 2 + 1))
            else:
                k = int(rng.choice(row)) + int(rng.integers(-1, 2))
            cuts.append(min(max(k, 0), len(q) - 1))
        _check(ctx, oracle, queries, choices, np.array(cuts, dtype=np.uint64), full)


def test_band_wide_windows(ctx, oracle):
    """Cutoffs up to 4094 (windows of up to 64 words, one pair per wave) on queries of 4,200-6,000
    units; the band kernel against the block kernel and the oracle."""
    queries, choices = _case(77, 3, 24, 4200, 6000, 4)
    full = oracle.cdist_mt(0, 0, queries, choices)
    for ks in ((4094, 4000, 3000), (2047, 2046, 1985), (640, 700, 1300)):
        cuts = np.array([min(k, len(q) - 1) for k, q in zip(ks, queries)], dtype=np.uint64)
        _check(ctx, oracle, queries, choices, cuts, full)


def test_band_many_choices(ctx, oracle):
    """More choices than one workgroup's chunk: several chunks per query, each wave scanning many
    64-choice batches and handing finished pairs' lanes to new candidates."""
    queries, choices = _case(91, 6, 2500, 100, 400, 4)
    full = oracle.cdist_mt(0, 0, queries, choices)
    for frac in (0.05, 0.2, 0.45):
        cuts = np.array([int(frac * len(q)) for q in queries], dtype=np.uint64)
        _check(ctx, oracle, queries, choices, cuts, full)


def test_band_single_cutoff_and_weights(ctx, oracle):
    """rfa_cdist with one cutoff for every query, and uniform weights (w, w, w): the band runs on
    ceil(cutoff / w) and the results scale by w, as in CachedLevenshtein with those weights."""
    import rapidfuzz_amd as rfa

    queries, choices = _case(5, 5, 120, 70, 600, 4)
    full = oracle.cdist_mt(0, 0, queries, choices)
    qs, cs = ctx.strings(queries, width=1), ctx.strings(choices, width=1)
    for w in (1, 3):
        sc = rfa.Scorer(ctx, 0, qs, weights=(w, w, w))
        for cu in (0, 5, 31, 62, 63, 64, 130, 300):
            got = sc.cdist(cs, 0, cutoff_u=cu)
            exp = oracle.cdist(0, 0, queries, choices, weights=(w, w, w), cutoff_u=cu)
            assert np.array_equal(got, exp), f"weights {w} cutoff {cu}: {np.count_nonzero(got != exp)} differ"


def test_band_autotune(ctx, oracle):
    """rfa_cdist_autotune measures words per lane x workgroup size per window-size bucket; the
    settings it keeps give the same results (and so do the rule's, before it runs)."""
    import rapidfuzz_amd as rfa

    queries, choices = _case(321, 8, 300, 120, 2600, 4)
    full = oracle.cdist_mt(0, 0, queries, choices)
    ks = [0, 3, 70, 150, 400, 700, 1300, 2000]
    cuts = np.array([min(k, len(q) - 1) for k, q in zip(ks, queries)], dtype=np.uint64)
    exp = np.where(full <= cuts[:, None], full, cuts[:, None] + 1).astype(np.uint64)
    sc = rfa.Scorer(ctx, 0, ctx.strings(queries, width=1))
    cs = ctx.strings(choices, width=1)
    assert np.array_equal(sc.cdist(cs, 0, cutoffs_u=cuts), exp)
    chosen = sc.cdist_autotune(cs, 0, cutoffs_u=cuts)
    used = [b for b in range(6) if chosen[b, 0]]
    assert len(used) >= 4, chosen
    for b in used:
        assert chosen[b, 0] in (1, 2) and chosen[b, 1] in (64, 128, 256), chosen
    assert np.array_equal(sc.cdist(cs, 0, cutoffs_u=cuts), exp)


def test_band_long_texts_small_cutoffs(ctx, oracle):
    """Long queries (2,600-5,200 units) under small cutoffs (1-6% of m): close matches (light
    mutations of the query) run every text unit, far past the point where unrelated strings and
    heavy mutations end, in waves whose other pairs come and go."""
    rng = np.random.default_rng(515)
    queries = _gen.random_strings(rng, 5, 2600, 5200, alpha_lo=65, alpha_n=4)
    choices = []
    for i in range(260):
        q = queries[i % 5]
        if i % 3 == 0:
            choices.append(_gen.mutate(rng, q, int(rng.integers(0, len(q) // 60 + 1)), alpha_lo=65, alpha_n=4))
        elif i % 3 == 1:
            choices.append(_gen.mutate(rng, q, int(rng.integers(len(q) // 40, len(q) // 8)), alpha_lo=65, alpha_n=4))
        else:
            n = max(1, len(q) + int(rng.integers(-60, 61)))
            choices.append(rng.integers(65, 69, size=n).astype(np.uint32))
    full = oracle.cdist_mt(0, 0, queries, choices)
    for frac in (0.01, 0.03, 0.06):
        cuts = np.array([int(frac * len(q)) for q in queries], dtype=np.uint64)
        exp = np.where(full <= cuts[:, None], full, cuts[:, None] + 1).astype(np.uint64)
        assert int(np.count_nonzero(full <= cuts[:, None])) >= 50
        got = _run(ctx, queries, choices, cuts, True)
        bad = np.argwhere(got != exp)
        assert len(bad) == 0, (frac, len(bad), [(int(i), int(j), int(got[i, j]), int(exp[i, j])) for i, j in bad[:4]])
    cuts = np.array([int(0.02 * len(q)) for q in queries], dtype=np.uint64)
    exp = np.where(full <= cuts[:, None], full, cuts[:, None] + 1).astype(np.uint64)
    assert np.array_equal(_run(ctx, queries, choices, cuts, False), exp)
