"""GPU parity of fuzz::WRatio (rfa_cdist, rfa_topk) against the oracle, bit for bit: the
ratio, the token ratio or the partial ratio and partial token ratio launched with the
cutoffs WRatio raises per pair, and the combining steps; short and long queries."""
import numpy as np
import pytest

import golden_cases as G
from test_gpu_token_sort import _sets

pytestmark = pytest.mark.gpu

WRATIO, SIM = 19, 1


def _check(ctx, oracle, q, c, width, signed8=False, cuts=(0.0, 50.0, 85.0, 95.0, 100.0, 100.5)):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, WRATIO, ctx.strings(q, width=width))
    if signed8:
        sc.set_option(rfa.SCORER_SIGNED_UNITS, 1)
    cs = ctx.strings(c, width=width)
    for cut in cuts:
        got = sc.cdist(cs, SIM, cutoff_d=cut)
        exp = oracle.cdist(WRATIO, SIM, q, c, cutoff_d=cut, char_width=width, signed8=signed8)
        bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
        assert len(bad) == 0, (width, signed8, cut, len(bad), [(int(i), int(j), got[i, j], exp[i, j])
                                                              for i, j in bad[:3]])
    return sc, cs


def test_known_answers(ctx):
    import rapidfuzz_amd as rfa

    n = 0
    for c in G.load()["cases"]:
        if c["metric"] != WRATIO:
            continue
        sc = rfa.Scorer(ctx, WRATIO, ctx.strings([G.expand(c["s1"])]))
        got = float(sc.cdist(ctx.strings([G.expand(c["s2"])]), SIM, cutoff_d=c.get("cutoff", 0.0))[0, 0])
        assert abs(got - c["expected"]) <= c["tol"], (G.case_id(c), got)
        n += 1
    assert n >= 7


@pytest.mark.parametrize("width,signed8", [(1, False), (1, True), (2, False), (4, False)])
def test_random_tokens(ctx, oracle, width, signed8):
    """Whitespace-rich strings of every unit width: choices up to about 9 times a query's
    length, so all three length-ratio branches occur, plus token permutations."""
    rng = np.random.default_rng(190 + width + 7 * signed8)
    q, c = _sets(rng, width, 48, 90, 24, 200)
    _check(ctx, oracle, q, c, width, signed8)


def test_sentences_and_branches(ctx, oracle):
    """Word sentences: equal, partial, reordered and much longer choices; launch cutoffs at
    and next to scores the pairs reach."""
    rng = np.random.default_rng(191)
    words = ["new", "york", "mets", "atlanta", "braves", "vs", "the", "wonderful", "city", "ny", "a", "hello"]

    def sent(n):
        return np.array([ord(ch) for ch in " ".join(words[int(i)] for i in rng.integers(0, len(words), size=n))],
                        np.uint32)

    q = [sent(int(rng.integers(1, 5))) for _ in range(70)]
    c = [sent(int(rng.integers(1, 16))) for _ in range(150)] + q[:20]
    c.append(np.array([ord(ch) for ch in "hello" + "abcde" * 7], np.uint32))
    c.append(np.zeros(0, np.uint32))
    base = oracle.cdist(WRATIO, SIM, q, c, cutoff_d=0.0, char_width=1)
    vals = np.unique(base[base > 0])
    cuts = [0.0] + [float(v) for v in vals[:: max(1, len(vals) // 6)]]
    cuts += [np.nextafter(x, 0.0) for x in cuts[1:3]] + [np.nextafter(x, 200.0) for x in cuts[1:3]]
    _check(ctx, oracle, q, c, 1, cuts=tuple(cuts))


@pytest.mark.parametrize("width", [1, 2])
def test_long_queries(ctx, oracle, width):
    """Queries of 65-900 units (the sub-scores' long paths: the block kernel's ratio and the
    long token-set kernel with per-pair cutoffs, the long partial_ratio path) against choices
    on both sides of 1.5 and 8 times their length, and a few short queries beside them."""
    from test_gpu_token_sort import _vocab_sentences

    rng = np.random.default_rng(195 + width)
    q, c = _vocab_sentences(rng, width, (15, 180), (1, 150), 10, 40)
    q = [x for x in q if 64 < len(x) <= 900] + [q[0][:20], q[1][:50]]
    c += [x[: len(x) // 3] for x in q[:4]] + [np.concatenate([x, x[:40]]) for x in q[:3]]
    _check(ctx, oracle, q, c, width, cuts=(0.0, 60.0, 90.0, 100.5))


def test_topk(ctx, oracle):
    """rfa_topk: the scores of the whole set, ordered by (score desc, index asc)."""
    import rapidfuzz_amd as rfa
    import _topk

    rng = np.random.default_rng(192)
    q, c = _sets(rng, 1, 20, 300, 24, 120)
    sc = rfa.Scorer(ctx, WRATIO, ctx.strings(q, width=1))
    cs = ctx.strings(c, width=1)
    for cut, k in ((0.0, 5), (60.0, 3), (90.0, 10)):
        idx, score = sc.topk(cs, k, kind=SIM, cutoff_d=cut)
        full = oracle.cdist(WRATIO, SIM, q, c, cutoff_d=cut, char_width=1)
        ei, es = _topk.expected_topk(full, cut, k)
        assert np.array_equal(idx, ei) and np.array_equal(score.view(np.uint64), es.view(np.uint64)), (cut, k)


def test_rejected(ctx):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, WRATIO, ctx.strings([b"new york"]))
    with pytest.raises(rfa.RfaError):
        sc.cdist(ctx.strings([b"york"]), rfa.DISTANCE)
    with pytest.raises(rfa.RfaError):
        sc.cdist(ctx.strings([b"york"]), SIM, semantics=rfa.MULTI)
