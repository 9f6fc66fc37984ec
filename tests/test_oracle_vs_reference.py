"""The oracle against the reference library itself (oracle/_ref/librfref.so,
compiled from the reference headers by oracle/Makefile), bit for bit.

This pins the oracle where the reference's known answers leave it open:
every metric x kind x cutoff, Cached* and experimental::Multi* semantics,
doubles at the normalization boundary (d / max = 1 - cutoff with the 1e-5
widening of common.hpp:50-53) and the Jaro-Winkler prefix pre-filter
(JaroWinkler_impl.hpp:54-62).  Skipped when the reference library is not built.
"""
import numpy as np
import pytest

import _gen

METRICS = [0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10]
MULTI_METRICS = [0, 1, 2, 3, 4, 5, 6]  # the reference has no Multi* for Hamming / Prefix / Postfix


def _kinds(metric):
    return [1] if metric in (6, 10) else [0, 1, 2, 3]


def _cutoffs(metric, kind):
    if metric in (0, 1, 2, 3, 7, 8, 9) and kind in (0, 1):
        return [(None, None), (0, None), (1, None), (2, None), (5, None), (17, None)]
    if metric in (6, 10):
        return [(None, 0.0), (None, 50.0), (None, 70.0), (None, 80.0), (None, 100.0)]
    return [(None, None), (None, 0.0), (None, 0.3), (None, 0.7), (None, 0.75), (None, 0.8), (None, 0.85),
            (None, 1.0)]


def _same(a, b):
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return np.array_equal(a, b)


def _strings(seed, nq=24, nc=40, qmax=90, cmax=140):
    rng = np.random.default_rng(seed)
    q = _gen.random_strings(rng, nq, 0, qmax, alpha_n=4)
    c = _gen.random_strings(rng, nc, 0, cmax, alpha_n=4)
    c += [_gen.mutate(rng, q[i], int(rng.integers(0, 5))) for i in range(nq)]
    # exact boundary ratios: lengths 5/10/20 with 1/2/4 edits (d / max = 0.2)
    for L in (5, 10, 20, 50):
        base = rng.integers(97, 101, size=L).astype(np.uint32)
        q.append(base)
        for e in range(0, 5):
            t = base.copy()
            t[:e] = 120  # e substitutions with a symbol absent from the base
            c.append(t)
    q += _gen.edge_strings()[:12]
    return q, c


@pytest.mark.parametrize("metric", METRICS)
def test_cached(oracle, reference, metric):
    # partial_ratio's oracle runs a naive DP per window: shorter strings
    q, c = _strings(100 + metric, qmax=40, cmax=70) if metric == 10 else _strings(100 + metric)
    weights_list = [(1, 1, 1)] + ([(1, 1, 2), (2, 2, 2), (3, 3, 7)] if metric == 0 else [])
    for weights in weights_list:
        for kind in _kinds(metric):
            for cu, cd in _cutoffs(metric, kind):
                exp = reference.cdist(metric, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd)
                got = oracle.cdist(metric, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd)
                if not _same(got, exp):
                    bad = np.argwhere((got.view(np.uint64) if got.dtype == np.float64 else got) !=
                                      (exp.view(np.uint64) if exp.dtype == np.float64 else exp))
                    i, j = bad[0]
                    raise AssertionError(f"metric {metric} w {weights} kind {kind} cutoff {cu}/{cd}: "
                                         f"{len(bad)} differ; q{i} c{j}: oracle {got[i, j]!r} reference {exp[i, j]!r}")


@pytest.mark.parametrize("metric", MULTI_METRICS)
@pytest.mark.parametrize("max_len", [8, 16, 32, 64])
def test_multi(oracle, reference, metric, max_len):
    rng = np.random.default_rng(7 * max_len + metric)
    q = _gen.random_strings(rng, 20, 0, max_len, alpha_n=4)
    c = _gen.random_strings(rng, 30, 0, 120, alpha_n=4)
    c += [_gen.mutate(rng, q[i], int(rng.integers(0, 4))) for i in range(10)]
    for kind in _kinds(metric):
        for cu, cd in _cutoffs(metric, kind):
            exp = reference.cdist(metric, kind, q, c, cutoff_u=cu, cutoff_d=cd, multi_len=max_len)
            got = oracle.cdist(metric, kind, q, c, cutoff_u=cu, cutoff_d=cd, multi=True)
            assert _same(got, exp), f"Multi metric {metric} MaxLen {max_len} kind {kind} cutoff {cu}/{cd}"


def test_jaro_winkler_prefix_filter(oracle, reference):
    """Common prefixes of 0-4 around the pre-filter cutoff (JaroWinkler_impl.hpp:54-62)."""
    rng = np.random.default_rng(3)
    q, c = [], []
    for _ in range(60):
        p = int(rng.integers(0, 5))
        pre = rng.integers(97, 101, size=p).astype(np.uint32)
        q.append(np.concatenate([pre, rng.integers(97, 105, size=int(rng.integers(1, 12))).astype(np.uint32)]))
        c.append(np.concatenate([pre, rng.integers(97, 105, size=int(rng.integers(1, 12))).astype(np.uint32)]))
    for pw in (0.1, 0.2, 0.25):
        for cd in (0.7, 0.71, 0.75, 0.8, 0.85, 0.9):
            for kind in (1, 3):
                exp = reference.cdist(5, kind, q, c, prefix_weight=pw, cutoff_d=cd)
                got = oracle.cdist(5, kind, q, c, prefix_weight=pw, cutoff_d=cd)
                assert _same(got, exp), f"JW pw {pw} cutoff {cd} kind {kind}"


@pytest.mark.parametrize("weights", [(1, 1, 0), (1, 1, 1), (1, 1, 2)])
def test_multi_levenshtein_weights(oracle, reference, weights):
    """MultiLevenshtein: uniform distance whatever the weights, maximum() from the weights."""
    rng = np.random.default_rng(sum(weights))
    q = _gen.random_strings(rng, 16, 0, 32, alpha_n=4)
    c = _gen.random_strings(rng, 30, 0, 60, alpha_n=4)
    for kind in (0, 1, 2, 3):
        for cu, cd in _cutoffs(0, kind):
            exp = reference.cdist(0, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd, multi_len=32)
            got = oracle.cdist(0, kind, q, c, weights=weights, cutoff_u=cu, cutoff_d=cd, multi=True)
            assert _same(got, exp), f"MultiLevenshtein w {weights} kind {kind} cutoff {cu}/{cd}"


def test_multi_levenshtein_rejects_weights(oracle):
    with pytest.raises(ValueError):
        oracle.cdist(0, 0, [[97]], [[98]], weights=(1, 1, 3), multi=True)


def test_positional_metrics_brute_force(oracle):
    """Hamming (padded), common prefix and common suffix against their definitions."""
    rng = np.random.default_rng(11)
    for _ in range(300):
        a = rng.integers(97, 100, size=int(rng.integers(0, 12))).astype(np.uint32)
        b = rng.integers(97, 100, size=int(rng.integers(0, 12))).astype(np.uint32)
        n = min(len(a), len(b))
        ham = max(len(a), len(b)) - int(np.count_nonzero(a[:n] == b[:n]))
        pre = next((i for i in range(n) if a[i] != b[i]), n)
        post = next((i for i in range(n) if a[len(a) - 1 - i] != b[len(b) - 1 - i]), n)
        assert oracle.score(7, 0, a, b) == ham
        assert oracle.score(8, 1, a, b) == pre
        assert oracle.score(9, 1, a, b) == post


def test_lcs_bits_equals_dp(oracle):
    """The oracle's bit-parallel LCS (the Indel scorers' raw value) against its plain DP:
    random strings over small and large alphabets, lengths across word boundaries."""
    import numpy as np

    rng = np.random.default_rng(77)
    for _ in range(400):
        alpha = int(rng.choice([2, 3, 26, 70000]))
        n1 = int(rng.choice([0, 1, 63, 64, 65, 127, 128, 129, int(rng.integers(0, 400))]))
        n2 = int(rng.integers(0, 400))
        a = rng.integers(0, alpha, size=n1).astype(np.uint32)
        b = rng.integers(0, alpha, size=n2).astype(np.uint32)
        assert oracle.lcs_bits(a, b) == oracle.lcs_raw(a, b), (n1, n2, alpha)
