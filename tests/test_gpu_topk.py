"""GPU parity of rfa_topk and the request server against the oracle's score
matrices, ordered by (score desc, index asc); scores compared bit for bit."""
import numpy as np
import pytest

import _gen
from _topk import expected_topk, vocabulary, word_strings

pytestmark = pytest.mark.gpu

LEV, INDEL, LCSSEQ, OSA, JARO, JW, RATIO = range(7)
SIM, NSIM = 1, 3


def _kind(metric):
    return SIM if metric in (JARO, JW, RATIO) else NSIM


def _assert_same(got, exp, what):
    gi, gs = got
    ei, es = exp
    bad = np.argwhere((gi != ei) | (gs.view(np.uint64) != es.view(np.uint64)))
    if bad.size:
        q = bad[0][0]
        raise AssertionError(f"{what}: {len(bad)} mismatching entries; query {q}: got {gi[q]} {gs[q]} "
                             f"expected {ei[q]} {es[q]}")


def _data(seed, nq=24, nc=1500):
    rng = np.random.default_rng(seed)
    vocab = vocabulary(rng)
    choices = word_strings(rng, nc, vocab, 1, 10)
    # long choices (> 128 units) and empty ones take the unrolled and ragged paths
    choices += [np.concatenate([choices[i], np.full(150, 32, np.uint32), choices[i]]) for i in range(5)]
    choices += [np.zeros(0, np.uint32)] * 3
    queries = [_gen.mutate(rng, choices[int(rng.integers(0, nc))], int(rng.integers(0, 6)), alpha_lo=97,
                           alpha_n=8) for _ in range(nq)]
    queries += [np.zeros(0, np.uint32), choices[0][:1]]
    queries += [np.concatenate([choices[1]] * 12)[:300]]  # > 256 units: dense path
    queries += [np.concatenate([choices[2]] * 8)[:200]]   # 4-word class
    queries += [np.concatenate([choices[3]] * 4)[:100]]   # 2-word class
    return queries, choices


@pytest.mark.parametrize("metric,weights,cutoff", [
    (RATIO, (1, 1, 1), 70.0), (RATIO, (1, 1, 1), 0.0), (RATIO, (1, 1, 1), 95.0),
    (INDEL, (1, 1, 1), 0.6), (LCSSEQ, (1, 1, 1), 0.5), (LEV, (1, 1, 1), 0.6), (LEV, (2, 2, 2), 0.7),
    (LEV, (1, 1, 2), 0.6), (OSA, (1, 1, 1), 0.6), (JW, (1, 1, 1), 0.8), (JARO, (1, 1, 1), 0.0)])
@pytest.mark.parametrize("k", [1, 5, 20])
def test_topk_vs_oracle(ctx, oracle, metric, weights, cutoff, k):
    import rapidfuzz_amd as rfa

    queries, choices = _data(7 + metric)
    kind = _kind(metric)
    scores = oracle.cdist(metric, kind, queries, choices, weights=weights, cutoff_d=cutoff)
    exp = expected_topk(scores, cutoff, k)
    sc = rfa.Scorer(ctx, metric, ctx.strings(queries, width=1), weights=weights)
    got = sc.topk(ctx.strings(choices, width=1), k, kind=kind, cutoff_d=cutoff)
    _assert_same(got, exp, f"metric {metric} w {weights} cutoff {cutoff} k {k}")


def test_topk_length_boundaries(ctx, oracle):
    """Choices one unit either side of the reachable-length limits, scores at the cutoff."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(5)
    queries, choices = [], []
    for m in (1, 5, 7, 10, 33, 64, 65, 100):
        q = rng.integers(97, 123, size=m).astype(np.uint32)
        queries.append(q)
        for extra in range(0, 3 * m + 3):
            choices.append(np.concatenate([q, rng.integers(48, 58, size=extra).astype(np.uint32)]))
        for cut in range(0, m + 1):
            choices.append(q[:cut].copy())
    for metric, cutoff in ((RATIO, 70.0), (RATIO, 50.0), (INDEL, 0.7), (LCSSEQ, 0.7), (LEV, 0.7), (LEV, 0.4)):
        kind = _kind(metric)
        scores = oracle.cdist(metric, kind, queries, choices, cutoff_d=cutoff)
        k = 400
        exp = expected_topk(scores, cutoff, k)
        got = rfa.Scorer(ctx, metric, ctx.strings(queries, width=1)).topk(ctx.strings(choices, width=1), k,
                                                                          kind=kind, cutoff_d=cutoff)
        _assert_same(got, exp, f"boundaries metric {metric} cutoff {cutoff}")


def test_topk_ties_and_overflow(ctx):
    """20,000 equal choices: more candidates than the per-query buffer (dense redo), ties by index."""
    import rapidfuzz_amd as rfa

    q = np.frombuffer(b"hello world", np.uint8).astype(np.uint32)
    choices = [q] * 20000 + [q[:5]]
    sc = rfa.Scorer(ctx, RATIO, ctx.strings([q, q[:5]], width=1))
    idx, s = sc.topk(ctx.strings(choices, width=1), 12, kind=SIM, cutoff_d=70.0)
    assert list(idx[0]) == list(range(12)) and np.all(s[0] == 100.0)
    assert idx[1][0] == 20000 and s[1][0] == 100.0


def test_topk_rejects_distance_kind(ctx):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, LEV, ctx.strings([b"abc"], width=1))
    with pytest.raises(rfa.InvalidArgument):
        sc.topk(ctx.strings([b"abd"], width=1), 1, kind=0)


def _server_case(seed=21, nq=60, nc=4000):
    rng = np.random.default_rng(seed)
    vocab = vocabulary(rng, 200)
    choices = word_strings(rng, nc, vocab, 1, 12)
    queries = [_gen.mutate(rng, choices[int(rng.integers(0, nc))], int(rng.integers(0, 7)), alpha_lo=97,
                           alpha_n=8) for _ in range(nq)]
    queries.append(np.concatenate([choices[0]] * 40)[:300])  # dense path inside the server
    return queries, choices


@pytest.mark.parametrize("metric,cutoff,cap", [(RATIO, 70.0, 65536), (RATIO, 70.0, 4), (INDEL, 0.5, 1024),
                                               (JW, 0.85, 1024)])
def test_server_alone_equals_mix(ctx, oracle, metric, cutoff, cap):
    """Each request's answer is the same served alone, in a burst, and under Poisson arrivals."""
    import rapidfuzz_amd as rfa

    queries, choices = _server_case()
    kind = _kind(metric)
    k = 5
    exp = expected_topk(oracle.cdist(metric, kind, queries, choices, cutoff_d=cutoff), cutoff, k)
    cs = ctx.strings(choices, width=1)
    srv = rfa.Server(ctx, cs, metric=metric, kind=kind, score_cutoff=cutoff, k=k, max_batch=16, slots=3,
                     candidate_cap=cap)
    try:
        alone = [srv.wait(srv.submit(q)) for q in queries]
        alone = (np.stack([a[0] for a in alone]), np.stack([a[1] for a in alone]))
        _assert_same(alone, exp, "alone")
        tickets = [srv.submit(q) for q in queries]
        burst = [srv.wait(t) for t in tickets]
        burst = (np.stack([b[0] for b in burst]), np.stack([b[1] for b in burst]))
        _assert_same(burst, exp, "burst")
        rng = np.random.default_rng(1)
        arrivals = np.cumsum(rng.exponential(1 / 2000.0, size=len(queries)))
        lat, idx, sc = srv.replay(queries, arrivals)
        _assert_same((idx, sc), exp, "replay")
        assert np.all(lat >= 0)
        st = srv.stats()
        assert st["requests"] == 3 * len(queries)
        assert st["max_batch_seen"] > 1  # the burst was batched
        if cap == 4:
            assert st["dense_fallbacks"] > 0
    finally:
        srv.close()
