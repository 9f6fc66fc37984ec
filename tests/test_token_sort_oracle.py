"""token_sort_ratio, partial_token_sort_ratio, QRatio, token_set_ratio and token_ratio:
pins of the oracle.

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

TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO, SIM = 11, 12, 13, 14, 15, 16, 17, 1
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


@pytest.mark.parametrize("metric", [TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO])
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


@pytest.mark.parametrize("metric", [TSR, PTSR, QR, TSET, TRATIO, PTSET, PTRATIO])
@pytest.mark.parametrize("s1,s2", [(b"South Korea", b"North Korea"), (b"bc", b"bca")])
def test_cutoff_boundary_issues(oracle, metric, s1, s2):
    # tests-fuzz.cpp:161-184
    score = float(oracle.score(metric, SIM, s1, s2, cutoff_d=0.0))
    assert float(oracle.score(metric, SIM, s1, s2, cutoff_d=score + 0.0001)) == 0.0
    assert abs(float(oracle.score(metric, SIM, s1, s2, cutoff_d=score - 0.0001)) - score) <= 1e-6


@pytest.mark.parametrize("metric", [TSET, TRATIO, PTSET, PTRATIO])
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
