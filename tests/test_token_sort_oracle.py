"""token_sort_ratio, partial_token_sort_ratio, QRatio, token_set_ratio, token_ratio and
WRatio: pins of the oracle.

* The reference's assertions (tests-fuzz.cpp) are in tests/golden/known_answers.json
  (metrics 11-13) and run in test_oracle_golden.
* detail::sorted_split(s).join() (common_impl.hpp:161-179): checked against what it
  must be -- the tokens of the string (maximal runs of non-whitespace), sorted, one
  space between -- for the whitespace set of each unit width and `char`'s signed
  order of 8-bit units.
* The scorers against the reference library for 8/16/32-bit units and `char`.
"""
import numpy as np
import pytest

TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO, WRATIO, SIM = 11, 12, 13, 14, 15, 16, 17, 19, 1
NARROW_WS = [0x09, 0x0A, 0x0B, 0x0C, 0x0D, 0x1C, 0x1D, 0x1E, 0x1F, 0x20]
WIDE_WS = NARROW_WS + [0x85, 0xA0, 0x1680] + list(range(0x2000, 0x200B)) + [0x2028, 0x2029, 0x202F, 0x205F, 0x3000]


def _expected_join(s, width, signed8):
    ws = set(NARROW_WS if width == 1 else WIDE_WS)
    toks, cur = [], []
    for c in s:
        if c in ws:
            if cur:
                toks.append(cur)
            cur = []
        else:
            cur.append(c)
    if cur:
        toks.append(cur)

    def key(t):
        return [((c ^ 0x80) - 0x80) if (width == 1 and signed8) else c for c in t]

    toks.sort(key=key)
    out = []
    for i, t in enumerate(toks):
        if i:
            out.append(0x20)
        out.extend(t)
    return out


def _random_text(rng, n, width):
    pool = [97, 98, 99, 100, 0x20, 0x20, 0x09, 0x0A, 0x1F, 0x85, 0xA0, 0xC3, 0xE9, 0xFF]
    if width > 1:
        pool += [0x3000, 0x2003, 0x1680, 0x4E2D, 0x0101]
    if width == 4:
        pool += [0x1F600]
    return np.array([pool[int(i)] for i in rng.integers(0, len(pool), size=n)], dtype=np.uint32)


@pytest.mark.parametrize("width,signed8", [(1, False), (1, True), (2, False), (4, False)])
def test_sorted_join(oracle, width, signed8):
    rng = np.random.default_rng(60 + width + 10 * signed8)
    for _ in range(400):
        s = _random_text(rng, int(rng.integers(0, 40)), width)
        got = [int(x) for x in oracle.sorted_join(s, char_width=width, signed8=signed8)]
        assert got == _expected_join([int(x) for x in s], width, signed8), (s, got)
        # idempotent: the join of a join is itself
        again = [int(x) for x in oracle.sorted_join(np.array(got, np.uint32), char_width=width, signed8=signed8)]
        assert again == got


def _sets(rng, width, nq=14, nc=18):
    q = [_random_text(rng, int(rng.integers(0, 30)), width) for _ in range(nq)]
    c = [_random_text(rng, int(rng.integers(0, 60)), width) for _ in range(nc)]
    # token permutations of the queries
    for s in q[:6]:
        toks = " ".join("".join(chr(x) for x in s).split(" ")[::-1])
        c.append(np.array([ord(ch) for ch in toks], np.uint32))
    return q, c


@pytest.mark.parametrize("metric", [TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO, WRATIO])
@pytest.mark.parametrize("width,signed8", [(1, False), (1, True), (2, False), (4, False)])
def test_vs_reference(oracle, reference, metric, width, signed8):
    rng = np.random.default_rng(70 + metric * 7 + width + 3 * signed8)
    q, c = _sets(rng, width)
    for cut in (0.0, 50.0, 80.0, 100.0):
        exp = reference.cdist(metric, SIM, q, c, cutoff_d=cut, char_width=-1 if signed8 else width)
        got = oracle.cdist(metric, SIM, q, c, cutoff_d=cut, char_width=width, signed8=signed8)
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), (metric, width, signed8, cut)


def test_signed_order_matters(oracle):
    # "\xe9t\xe9 abc": as `char`, 0xE9 < 'a' (negative), so the accented token sorts first
    s = np.array([0xE9, 0x74, 0xE9, 0x20, 0x61, 0x62, 0x63], np.uint32)
    assert [int(x) for x in oracle.sorted_join(s, char_width=1, signed8=True)] == \
        [0xE9, 0x74, 0xE9, 0x20, 0x61, 0x62, 0x63]
    assert [int(x) for x in oracle.sorted_join(s, char_width=1, signed8=False)] == \
        [0x61, 0x62, 0x63, 0x20, 0xE9, 0x74, 0xE9]


@pytest.mark.parametrize("metric", [TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO, WRATIO])
@pytest.mark.parametrize("s1,s2", [(b"South Korea", b"North Korea"), (b"bc", b"bca")])
def test_cutoff_boundary_issues(oracle, metric, s1, s2):
    # tests-fuzz.cpp:161-184
    score = float(oracle.score(metric, SIM, s1, s2, cutoff_d=0.0))
    assert float(oracle.score(metric, SIM, s1, s2, cutoff_d=score + 0.0001)) == 0.0
    assert abs(float(oracle.score(metric, SIM, s1, s2, cutoff_d=score - 0.0001)) - score) <= 1e-6


@pytest.mark.parametrize("metric", [TSET, TRATIO, PTSET, PTRATIO, WRATIO])
def test_disjoint_token_sets_with_repeats(oracle, reference, metric):
    """No token in common but shared characters, and repeated tokens on each side: the
    paths where partial_token_ratio's second partial_ratio (deduplicated differences)
    and token_set_ratio's difference LCS decide the score."""
    rng = np.random.default_rng(80 + metric)
    qv, cv = ["ab", "ba", "aab", "bba", "abab"], ["a", "b", "abb", "bab", "baa", "aaab"]

    def sent(vocab, n):
        return np.array([ord(ch) for ch in " ".join(vocab[int(i)] for i in rng.integers(0, len(vocab), size=n))],
                        np.uint32)

    q = [sent(qv, int(rng.integers(1, 7))) for _ in range(14)]
    c = [sent(cv, int(rng.integers(1, 9))) for _ in range(16)]
    for cut in (0.0, 40.0, 75.0):
        exp = reference.cdist(metric, SIM, q, c, cutoff_d=cut, char_width=1)
        got = oracle.cdist(metric, SIM, q, c, cutoff_d=cut, char_width=1)
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), (metric, cut)


def _wratio_by_hand(oracle, s1, s2, cut):
    """WRatio written out from its sub-scores (fuzz_impl.hpp:838-876), each taken from the
    oracle with the raised cutoff: a check of the oracle's combining steps."""
    if cut > 100 or not len(s1) or not len(s2):
        return 0.0
    l1, l2 = len(s1), len(s2)
    lr = l1 / l2 if l1 > l2 else l2 / l1
    end = float(oracle.score(6, SIM, s1, s2, cutoff_d=cut))
    if lr < 1.5:
        c = max(cut, end) / 0.95
        return max(end, float(oracle.score(TRATIO, SIM, s1, s2, cutoff_d=c)) * 0.95)
    ps = 0.9 if lr <= 8.0 else 0.6
    c = max(cut, end) / ps
    end = max(end, float(oracle.score(10, SIM, s1, s2, cutoff_d=c)) * ps)
    c = max(c, end) / 0.95
    return max(end, float(oracle.score(PTRATIO, SIM, s1, s2, cutoff_d=c)) * 0.95 * ps)


def test_wratio_branches(oracle, reference):
    """Every length-ratio branch (< 1.5, 1.5 - 8, > 8), cutoffs at the sub-scores' raised
    cutoffs and next to them, against the reference library and the formula written out."""
    rng = np.random.default_rng(19)
    words = ["new", "york", "mets", "atlanta", "braves", "vs", "the", "city", "ny", "a", "b", "ab"]

    def sent(n):
        return " ".join(words[int(i)] for i in rng.integers(0, len(words), size=n)).encode()

    q = [sent(int(rng.integers(1, 4))) for _ in range(10)]
    c = [sent(int(rng.integers(1, 12))) for _ in range(24)] + [b"x" * 70, b"new york" * 9, b"mets"]
    qa = [np.frombuffer(x, np.uint8).astype(np.uint32) for x in q]
    ca = [np.frombuffer(x, np.uint8).astype(np.uint32) for x in c]
    base = oracle.cdist(WRATIO, SIM, qa, ca, cutoff_d=0.0, char_width=1)
    for cut in (0.0, 30.0, 60.0, 85.0, 90.0, 95.0, 100.0, 100.5):
        exp = reference.cdist(WRATIO, SIM, qa, ca, cutoff_d=cut, char_width=1)
        got = oracle.cdist(WRATIO, SIM, qa, ca, cutoff_d=cut, char_width=1)
        assert np.array_equal(got.view(np.uint64), exp.view(np.uint64)), cut
    for i, s1 in enumerate(qa):
        for j, s2 in enumerate(ca):
            v = float(base[i, j])
            for cut in (0.0, v, v + 1e-9, max(0.0, v - 1e-9), 50.0):
                got = float(oracle.score(WRATIO, SIM, s1, s2, cutoff_d=cut))
                assert got == _wratio_by_hand(oracle, s1, s2, cut), (q[i], c[j], cut)
