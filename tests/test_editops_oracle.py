"""The oracle's editops (levenshtein_editops, Levenshtein_impl.hpp:959-1015 and
1229-1257; lcs_seq_editops = indel_editops, LCSseq_impl.hpp:440-534; hamming_editops,
Hamming_impl.hpp:40-59), pinned:

* against the compiled reference, operation for operation: Levenshtein pairs whose
  first string has at most 64 units after the common affix is removed (where the
  reference walks the full single-word matrix) and longer ones up to 4096 units below
  its Hirschberg split (where it walks the band of its block recurrence; the oracle
  restates that band, and its walk is also checked against the full-matrix walk),
  LCS pairs up to 4096 units (its single-word, unrolled and blockwise matrices);
* by what any correct script satisfies: applied to s1 it gives s2 (the reference's
  own editops tests, tests-Levenshtein.cpp:235-316 and tests-Indel.cpp:275-285, are
  of this form and their strings are used here), its length is the distance, and
  positions never decrease.
"""
import numpy as np
import pytest

import _gen

LEV, INDEL, LCS, HAMMING = 0, 1, 2, 7
REPLACE, INSERT, DELETE = 1, 2, 3


def apply_ops(ops, s1, s2):
    """editops_apply (types.hpp) restated: copy s1, applying each op at its positions."""
    t, sp, dp = ops
    out = []
    src = 0
    for k in range(len(t)):
        while src < sp[k]:
            out.append(int(s1[src]))
            src += 1
        if t[k] == DELETE:
            src += 1
        elif t[k] == INSERT:
            out.append(int(s2[dp[k]]))
        elif t[k] == REPLACE:
            out.append(int(s2[dp[k]]))
            src += 1
    while src < len(s1):
        out.append(int(s1[src]))
        src += 1
    return out


def _codes(s):
    return np.frombuffer(s.encode(), np.uint8).astype(np.uint32) if isinstance(s, str) else s


REF_STRINGS = [
    ("Lorem ipsum.", "XYZLorem ABC iPsum"),
    ("b", "aaaaaaaaaaaaaaaabbaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa"),
    ("aa", "abb"),
]


@pytest.mark.parametrize("metric", [LEV, LCS, HAMMING])
def test_apply_and_count(oracle, metric):
    rng = np.random.default_rng(5 + metric)
    pairs = [(_codes(a), _codes(b)) for a, b in REF_STRINGS]
    pairs += [(_gen.random_strings(rng, 1, 0, 40, alpha_n=3)[0], _gen.random_strings(rng, 1, 0, 50, alpha_n=3)[0])
              for _ in range(300)]
    for s1, s2 in pairs:
        ops = oracle.editops(metric, s1, s2)
        assert apply_ops(ops, s1, s2) == [int(x) for x in s2]
        kind_metric = metric if metric in (LEV, HAMMING) else INDEL
        assert len(ops[0]) == int(oracle.score(kind_metric, 0, s1, s2))
        assert np.all(np.diff(ops[1].astype(np.int64)) >= 0) and np.all(np.diff(ops[2].astype(np.int64)) >= 0)


def _check_vs_reference(oracle, reference, metric, pairs):
    for s1, s2 in pairs:
        got = oracle.editops(metric, s1, s2)
        exp = reference.editops(metric, s1, s2)
        for g, e in zip(got, exp):
            assert np.array_equal(g, e), f"metric {metric} {list(s1)} {list(s2)}: {got} vs {exp}"


def test_levenshtein_vs_reference(oracle, reference):
    rng = np.random.default_rng(21)
    pairs = [(_codes(a), _codes(b)) for a, b in REF_STRINGS]
    for alpha in (2, 3, 6):
        for _ in range(150):
            s1 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 65))).astype(np.uint32)
            s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 150))).astype(np.uint32)
            pairs.append((s1, s2))
            pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(0, 8)), alpha_n=alpha)))
    _check_vs_reference(oracle, reference, LEV, pairs)


def _long_lev_pairs(rng, count):
    out = []
    for it in range(count):
        alpha = int(rng.choice([2, 3, 4, 26]))
        n1 = int(rng.choice([int(rng.integers(65, 200)), int(rng.integers(200, 1500)), int(rng.integers(1500, 4097))]))
        s1 = rng.integers(97, 97 + alpha, size=n1).astype(np.uint32)
        kind = it % 4
        if kind == 0:
            s2 = _gen.mutate(rng, s1, int(rng.integers(0, max(1, n1 // 3))), alpha_n=alpha)
        elif kind == 1:
            s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(1, 1000))).astype(np.uint32)
        elif kind == 2:
            pre = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 300))).astype(np.uint32)
            s2 = np.concatenate([pre, _gen.mutate(rng, s1, 5, alpha_n=alpha)])
        else:  # a second string under 10 units: never split
            s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(1, 10))).astype(np.uint32)
        if (2 * len(s1) * len(s2)) // 8 < 1048576 or len(s2) < 10:  # below the Hirschberg split
            out.append((s1, s2))
    return out


def test_levenshtein_long_vs_reference(oracle, reference):
    """Stripped first strings of 65-4096 units: the block band restated, against the reference,
    and the full-matrix walk (independent of any band) giving the same script."""
    rng = np.random.default_rng(23)
    pairs = _long_lev_pairs(rng, 160)
    assert len(pairs) > 100
    _check_vs_reference(oracle, reference, LEV, pairs)
    oracle.set_lev_block(False)
    try:
        for s1, s2 in pairs[:60]:
            full = oracle.editops(LEV, s1, s2)
            oracle.set_lev_block(True)
            band = oracle.editops(LEV, s1, s2)
            oracle.set_lev_block(False)
            assert all(np.array_equal(x, y) for x, y in zip(full, band))
    finally:
        oracle.set_lev_block(True)


def test_levenshtein_hirschberg_vs_reference(oracle, reference):
    """Pairs the reference splits by Hirschberg (stripped |s1| > 64, |s2| >= 10 and
    |s1| * |s2| >= 2^22): the oracle's split at the first smallest sum of exact forward and
    backward rows, its leaves walked over full tables, against the reference."""
    rng = np.random.default_rng(25)
    pairs = []
    for it in range(24):
        alpha = int(rng.choice([2, 4, 26]))
        n1 = int(rng.integers(2050, 3600))
        s1 = rng.integers(97, 97 + alpha, size=n1).astype(np.uint32)
        kind = it % 3
        if kind == 0:
            s2 = _gen.mutate(rng, s1, int(rng.integers(1, n1 // 4)), alpha_n=alpha)
        elif kind == 1:
            s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(2050, 3600))).astype(np.uint32)
        else:
            pre = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 500))).astype(np.uint32)
            s2 = np.concatenate([pre, _gen.mutate(rng, s1, 30, alpha_n=alpha)])
        pairs.append((s1, s2))
    _check_vs_reference(oracle, reference, LEV, pairs)


def test_hamming_vs_reference(oracle, reference):
    """hamming_editops with pad (Hamming_impl.hpp:40-59), tests-Hamming.cpp:112-127's strings included."""
    rng = np.random.default_rng(23)
    pairs = [(_codes("Lorem ipsum."), _codes("XYZLorem ABC iPsum")), (_codes("XYZLorem ABC iPsum"),
                                                                      _codes("Lorem ipsum.")),
             (_codes(""), _codes("ab"))]
    for _ in range(200):
        s1 = rng.integers(97, 100, size=int(rng.integers(0, 80))).astype(np.uint32)
        s2 = rng.integers(97, 100, size=int(rng.integers(0, 80))).astype(np.uint32)
        pairs.append((s1, s2))
    _check_vs_reference(oracle, reference, HAMMING, pairs)


def test_lcs_vs_reference(oracle, reference):
    rng = np.random.default_rng(22)
    pairs = [(_codes(a), _codes(b)) for a, b in REF_STRINGS]
    for alpha in (2, 4, 8):
        for _ in range(120):
            s1 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 200))).astype(np.uint32)
            s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 200))).astype(np.uint32)
            pairs.append((s1, s2))
            pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(0, 8)), alpha_n=alpha)))
    # several words, through the reference's unrolled (<= 8 words) and blockwise LCS paths
    for n1 in (65, 300, 520, 1500, 3000, 4096):
        s1 = rng.integers(97, 101, size=n1).astype(np.uint32)
        pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(1, 40)))))
        pairs.append((s1, rng.integers(97, 101, size=int(rng.integers(0, 300))).astype(np.uint32)))
    # tests-Indel.cpp:275-285
    pairs.append((_codes("bbbdbbmbbbbbbbbbBbfbbbbbbbbbbbbbbbbbbbrbbbbbrbbbbbdbnbbbjbhbbbbbbbbbhbbbbbCbobb"
                         "bxbbbbbkbbbAbxbbwbbbtbcbbbbebbiblbbbbqbbbbbbpbbbbbbubbbkbbDbbbhbkbCbbgbbrbbbbbb"
                         "bbbbbkbyvbbsbAbbbbz"),
                  _codes("jaaagaaqyaaaanrCfwaaxaeahtaaaCzaaaspaaBkvaaaaqDaacndaaeolwiaaauaaaaaaamA")))
    _check_vs_reference(oracle, reference, LCS, pairs)
    for s1, s2 in pairs[-1:]:
        assert len(oracle.editops(LCS, s1, s2)[0]) == 231
