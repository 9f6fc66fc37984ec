"""The oracle's fuzz::partial_ratio_alignment (score and the source / destination windows,
fuzz_impl.hpp:68-230) on the CPU: the reference's own known answer and the compiled
reference on random pairs -- small alphabets (ties between windows), both length orders,
equal lengths (the swapped second pass), empty strings, cutoffs at and next to the scores
the pairs reach."""
import numpy as np
import pytest

import _gen


def _u(s):
    return [ord(ch) for ch in s]


def test_known_answer_issue_231(oracle):
    # tests-fuzz.cpp:187-200
    s1 = ("er merkantilismus f/rderte handel und verkehr mit teils marktkonformen, teils "
          "dirigistischen ma_nahmen.")
    s2 = ("ils marktkonformen, teils dirigistischen ma_nahmen. an der schwelle zum 19. "
          "jahrhundert entstand ein neu")
    score, ss, se, ds, de = oracle.partial_ratio_alignment(_u(s1), _u(s2), 0.0)
    assert abs(score - 66.2337662) < 1e-6
    assert (ss, se, ds, de) == (0, 103, 0, 51)


def _pairs(rng, count):
    out = []
    for it in range(count):
        alpha = int(rng.choice([2, 3, 4, 26]))
        l1 = int(rng.integers(0, 40))
        kind = it % 4
        if kind == 0:  # needle inside a longer haystack, lightly mutated
            a = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
            pre = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 30))).astype(np.uint32)
            post = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 30))).astype(np.uint32)
            b = np.concatenate([pre, _gen.mutate(rng, a, int(rng.integers(0, 3)), alpha_n=alpha), post])
        elif kind == 1:  # equal lengths
            a = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
            b = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
        else:  # independent lengths, either order
            a = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
            b = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 70))).astype(np.uint32)
        if it % 7 == 3:
            a, b = b, a
        out.append((a, b))
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_alignment_equals_reference(oracle, reference, seed):
    rng = np.random.default_rng(300 + seed)
    n = 0
    for a, b in _pairs(rng, 400):
        base = reference.partial_ratio_alignment(a, b, 0.0)
        cuts = [0.0, 50.0, 100.0, 100.5, base[0], np.nextafter(base[0], 200.0), np.nextafter(base[0], 0.0)]
        for c in cuts:
            exp = reference.partial_ratio_alignment(a, b, c)
            got = oracle.partial_ratio_alignment(a, b, c)
            assert got == exp, (a.tolist(), b.tolist(), c, got, exp)
            n += 1
    assert n >= 2800


def test_alignment_windows_score(oracle):
    """The reported windows reproduce the score: ratio(s1[src], s2[dest]) equals it whenever
    the score is positive (a property of the result, independent of the search order)."""
    rng = np.random.default_rng(310)
    checked = 0
    for a, b in _pairs(rng, 300):
        score, ss, se, ds, de = oracle.partial_ratio_alignment(a, b, 0.0)
        if score <= 0 or len(a) == 0 or len(b) == 0:
            continue
        x, y = a[ss:se], b[ds:de]
        lcs = oracle.lcs_raw(x, y)
        tot = len(x) + len(y)
        r = 100.0 * (1.0 - (tot - 2 * lcs) / tot) if tot else 100.0
        assert abs(r - score) < 1e-9, (a.tolist(), b.tolist(), score, (ss, se, ds, de), r)
        checked += 1
    assert checked > 100
