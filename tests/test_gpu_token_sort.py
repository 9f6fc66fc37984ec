"""GPU parity of token_sort_ratio, partial_token_sort_ratio, QRatio (rfa_cdist / rfa_topk),
token_set_ratio, token_ratio, partial_token_set_ratio and partial_token_ratio (rfa_cdist)
against the oracle, bit for bit: the device
token sort (plain and deduplicated) for 8/16/32-bit units and signed 8-bit order, the
cached token-sorted choice sets, the masked-LCS set kernel and its long-query form."""
import numpy as np
import pytest

import golden_cases as G

pytestmark = pytest.mark.gpu

TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO, SIM = 11, 12, 13, 14, 15, 16, 17, 1


def _text(rng, n, width):
    pool = [97, 98, 99, 100, 101, 0x20, 0x20, 0x20, 0x09, 0x0A, 0x1F, 0x85, 0xA0, 0xC3, 0xE9, 0xFF]
    if width > 1:
        pool += [0x3000, 0x2003, 0x1680, 0x4E2D, 0x0101]
    if width == 4:
        pool += [0x1F600]
    return np.array([pool[int(i)] for i in rng.integers(0, len(pool), size=n)], dtype=np.uint32)


def _sets(rng, width, nq, nc, qmax, cmax):
    q = [_text(rng, int(rng.integers(0, qmax)), width) for _ in range(nq)]
    c = [_text(rng, int(rng.integers(0, cmax)), width) for _ in range(nc)]
    for s in q[:20]:  # token permutations of queries
        toks = "".join(chr(x) for x in s).split(" ")
        rng.shuffle(toks)
        c.append(np.array([ord(ch) for ch in " ".join(toks)], np.uint32))
    return q, c


def _check(ctx, oracle, metric, q, c, width, signed8=False, cuts=(0.0, 60.0, 90.0)):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, metric, ctx.strings(q, width=width))
    if signed8:
        sc.set_option(rfa.SCORER_SIGNED_UNITS, 1)
    cs = ctx.strings(c, width=width)
    for cut in cuts:
        got = sc.cdist(cs, SIM, cutoff_d=cut)
        exp = oracle.cdist(metric, SIM, q, c, cutoff_d=cut, char_width=width, signed8=signed8)
        bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
        assert len(bad) == 0, (metric, width, signed8, cut, len(bad), bad[:3].tolist())
    return sc, cs


def test_known_answers(ctx):
    import rapidfuzz_amd as rfa

    for c in G.load()["cases"]:
        if c["metric"] not in (TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO):
            continue
        sc = rfa.Scorer(ctx, c["metric"], ctx.strings([G.expand(c["s1"])]))
        got = float(sc.cdist(ctx.strings([G.expand(c["s2"])]), SIM, cutoff_d=c.get("cutoff", 0.0))[0, 0])
        assert abs(got - c["expected"]) <= c["tol"], (G.case_id(c), got)


@pytest.mark.parametrize("metric", [TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO])
@pytest.mark.parametrize("width,signed8", [(1, False), (1, True), (2, False), (4, False)])
def test_short(ctx, oracle, metric, width, signed8):
    """Queries of 0-40 units (tile path) against choices of 0-90 units."""
    rng = np.random.default_rng(500 + metric * 11 + width + 5 * signed8)
    q, c = _sets(rng, width, 80, 60, 40, 90)
    _check(ctx, oracle, metric, q, c, width, signed8)


@pytest.mark.parametrize("metric", [TSR, PTSR])
def test_long(ctx, oracle, metric):
    """Queries of 70-200 units (block / long partial paths) and many tokens per string."""
    rng = np.random.default_rng(560 + metric)
    q, c = _sets(rng, 1, 6, 12, 200, 260)
    q = [s for s in q if len(s) > 64] + [_text(rng, 150, 1)]
    _check(ctx, oracle, metric, q, c, 1, cuts=(0.0, 70.0))


def test_cached_choices_and_option_switch(ctx, oracle):
    """A second choice set, the first one again, then the signed order: every call
    scores the right token-sorted copy."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(590)
    q, c1 = _sets(rng, 1, 30, 40, 30, 60)
    _, c2 = _sets(rng, 1, 30, 40, 30, 60)
    sc = rfa.Scorer(ctx, TSR, ctx.strings(q, width=1))
    s1, s2 = ctx.strings(c1, width=1), ctx.strings(c2, width=1)
    for chs, cs in ((c1, s1), (c2, s2), (c1, s1)):
        got = sc.cdist(cs, SIM)
        assert np.array_equal(got.view(np.uint64), oracle.cdist(TSR, SIM, q, chs, char_width=1).view(np.uint64))
    sc.set_option(rfa.SCORER_SIGNED_UNITS, 1)
    got = sc.cdist(s1, SIM)
    exp = oracle.cdist(TSR, SIM, q, c1, char_width=1, signed8=True)
    assert np.array_equal(got.view(np.uint64), exp.view(np.uint64))


@pytest.mark.parametrize("metric", [TSR, QR, TSET, TRATIO, PTSET, PTRATIO])
def test_topk(ctx, oracle, metric):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(600 + metric)
    q, c = _sets(rng, 1, 8, 400, 25, 60)
    sc = rfa.Scorer(ctx, metric, ctx.strings(q, width=1))
    idx, score = sc.topk(ctx.strings(c, width=1), 5, cutoff_d=50.0)
    full = oracle.cdist(metric, SIM, q, c, cutoff_d=50.0, char_width=1)
    for i in range(len(q)):
        order = sorted(((-full[i, j], j) for j in range(len(c)) if full[i, j] >= 50.0))[:5]
        assert list(idx[i]) == [j for _, j in order] + [-1] * (5 - len(order))
        assert np.array_equal(score[i], np.array([-s for s, _ in order] + [0.0] * (5 - len(order))))


def test_token_sets_many_choice_tokens(ctx, oracle):
    """Choices of up to 400 units (more than 64 distinct tokens: the set kernel recomputes
    membership past its 64-token mask), shared tokens, 32- and 64-bit query tiles."""
    rng = np.random.default_rng(620)
    words = [np.array([ord(ch) for ch in w], np.uint32) for w in
             ["a", "b", "ab", "ba", "abc", "cab", "x", "yy", "zz", "q", "r", "st", "uv", "w", "aa", "bb"]]
    def sentence(n):
        parts = [words[int(i)] for i in rng.integers(0, len(words), size=n)]
        out = []
        for i, p in enumerate(parts):
            if i:
                out.append(32)
            out.extend(int(x) for x in p)
        return np.array(out, np.uint32)
    q = [sentence(int(rng.integers(0, 12))) for _ in range(70)]
    q = [x[:64] for x in q]
    c = [sentence(int(rng.integers(0, 20))) for _ in range(50)]
    c += [np.concatenate([rng.integers(200, 255, size=300).astype(np.uint32) * 0 + 32 +
                          (np.arange(300, dtype=np.uint32) % 2) * (np.arange(300, dtype=np.uint32) + 1) % 200,
                          sentence(5)]) for _ in range(4)]
    for metric in (TSET, TRATIO, PTSET, PTRATIO):
        _check(ctx, oracle, metric, q, c, 1, cuts=(0.0, 70.0, 100.5))


@pytest.mark.parametrize("signed8", [False, True])
def test_partial_token_ratio_few_repeats(ctx, oracle, signed8):
    """partial_token_ratio where few strings repeat a token: the deduplicated joins' partial_ratio
    is computed only for the rows and columns of the strings that repeat one (about 10% of the
    choices, 2 of 40 queries), the rest taken from the sorted joins'."""
    rng = np.random.default_rng(640 + signed8)
    vocab = [np.array([int(x) for x in rng.integers(97, 123, size=int(rng.integers(2, 6)))], np.uint32)
             for _ in range(400)]
    if signed8:
        vocab += [np.array([0xE9, 97 + i % 3], np.uint32) for i in range(20)]
    def sentence(n, repeat):
        idx = list(rng.choice(len(vocab), size=n, replace=False))
        if repeat and n:
            idx.append(idx[int(rng.integers(0, len(idx)))])
            rng.shuffle(idx)
        out = []
        for i, t in enumerate(idx):
            if i:
                out.append(32)
            out.extend(int(x) for x in vocab[int(t)])
        return np.array(out, np.uint32)
    q = [sentence(int(rng.integers(1, 5)), i in (5, 17))[:64] for i in range(40)]
    c = [sentence(int(rng.integers(0, 12)), rng.random() < 0.1) for _ in range(300)]
    c += [q[i] for i in range(0, 40, 4)]
    _check(ctx, oracle, PTRATIO, q, c, 1, signed8=signed8, cuts=(0.0, 60.0, 95.0))


def _vocab_sentences(rng, width, nwords_q, nwords_c, nq, nc, base=97):
    """Sentences over a vocabulary of random words: long deduplicated joins, shared tokens."""
    hi = base + 6
    vocab = [rng.integers(base, hi, size=int(rng.integers(1, 7))).astype(np.uint32) for _ in range(900)]
    if width > 1:
        vocab += [np.array([0x4E2D + i % 5, 0x0101], np.uint32) for i in range(40)]

    def sentence(n):
        idx = rng.integers(0, len(vocab), size=n)
        out = []
        for i, t in enumerate(idx):
            if i:
                out.append(32 if i % 7 else 0x09)
            out.extend(int(x) for x in vocab[int(t)])
        return np.array(out, np.uint32)
    q = [sentence(int(rng.integers(*nwords_q))) for _ in range(nq)]
    c = [sentence(int(rng.integers(*nwords_c))) for _ in range(nc)]
    # choices sharing many of a query's tokens, in another order, plus some of their own
    for s in q[:6]:
        toks = "".join(chr(x) for x in s).split()
        rng.shuffle(toks)
        keep = toks[: max(1, len(toks) * 2 // 3)] + ["zz", "qq"]
        c.append(np.array([ord(ch) for ch in " ".join(keep)], np.uint32))
    c += [q[1], np.zeros(0, np.uint32)]
    return q, c


@pytest.mark.parametrize("width", [1, 2, 4])
@pytest.mark.parametrize("metric", [TSET, TRATIO, PTSET, PTRATIO])
def test_token_sets_long_queries(ctx, oracle, metric, width):
    """Queries whose deduplicated join has 65-1000 units (the long set kernel: the join as a
    pattern of several words across the lanes), with short queries in the same scorer."""
    rng = np.random.default_rng(700 + 7 * metric + width)
    q, c = _vocab_sentences(rng, width, (20, 250), (0, 60), 10, 40)
    q = [x for x in q if 64 < len(x) <= 1000] + [q[0][:30]]
    _check(ctx, oracle, metric, q, c, width, cuts=(0.0, 70.0, 100.5))


def test_token_set_ratio_very_long_queries(ctx, oracle):
    """token_set_ratio / token_ratio with deduplicated joins of 1,000-4,000 units (up to 64
    pattern words), against choices of up to 2,000 units."""
    rng = np.random.default_rng(720)
    q, c = _vocab_sentences(rng, 1, (300, 900), (0, 500), 5, 20)
    q = [x for x in q if len(x) <= 4096]
    for metric in (TSET, TRATIO):
        _check(ctx, oracle, metric, q, c, 1, cuts=(0.0, 80.0))


def test_token_set_limits(ctx):
    import rapidfuzz_amd as rfa

    words = " ".join(f"w{i:04d}" for i in range(900)).encode()  # a join of 5,399 units
    long_q = rfa.Scorer(ctx, TSET, ctx.strings([words]))
    with pytest.raises(rfa.RfaError):
        long_q.cdist(ctx.strings([b"w0001"]), SIM)  # deduplicated join over 4096 units
    sc = rfa.Scorer(ctx, TRATIO, ctx.strings([b"a b"]))
    with pytest.raises(rfa.InvalidArgument):
        sc.topk(ctx.strings([b"b a"]), 1, kind=3)  # similarity only


def test_rejected(ctx):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, TSR, ctx.strings([b"a b"]))
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"b a"]), 0)
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"b a"]), SIM, semantics=1)
    with pytest.raises(rfa.InvalidArgument):
        sc.set_option(99, 1)


def test_partial_token_ratio_signed_switch(ctx, oracle):
    """partial_token_ratio's per-scorer caches (the repeat rows and their scorer) follow a switch
    of RFA_SCORER_SIGNED_UNITS: queries and choices that repeat tokens with units >= 0x80,
    scored unsigned, then signed, then unsigned again on one scorer."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(597)
    toks = [[0xE9, 0x61], [0x61, 0xE9], [0x62], [0xC3, 0x62, 0x62], [0x7A]]

    def sent(n):
        out = []
        for i in rng.integers(0, len(toks), size=n):
            if out:
                out.append(0x20)
            out.extend(toks[int(i)])
        return np.array(out, np.uint32)

    q = [sent(int(rng.integers(1, 6))) for _ in range(40)]
    c = [sent(int(rng.integers(1, 9))) for _ in range(70)]
    sc = rfa.Scorer(ctx, PTRATIO, ctx.strings(q, width=1))
    cs = ctx.strings(c, width=1)
    for signed8 in (False, True, False):
        sc.set_option(rfa.SCORER_SIGNED_UNITS, 1 if signed8 else 0)
        for cut in (0.0, 70.0):
            got = sc.cdist(cs, SIM, cutoff_d=cut)
            exp = oracle.cdist(PTRATIO, SIM, q, c, cutoff_d=cut, char_width=1, signed8=signed8)
            assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), (signed8, cut)
