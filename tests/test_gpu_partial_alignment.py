"""GPU parity of rfa_partial_ratio_alignment (fuzz::partial_ratio_alignment, the score and
the source / destination windows) against the oracle, field for field and bit for bit:
the reference's issue-231 known answer, random pairs over small alphabets (ties between
windows), both length orders, equal lengths (the swapped second pass), empty strings,
needles of 1, 2-4 and 5-16 words and past 1024 units up to the 4096-unit limit,
8/16/32-bit units, and cutoffs at and next to the scores the pairs reach."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu


def _u(s):
    return np.array([ord(ch) for ch in s], dtype=np.uint32)


def _small_pairs(rng, count, alpha_max=26):
    out = []
    for it in range(count):
        alpha = int(rng.choice([2, 3, 4, alpha_max]))
        l1 = int(rng.integers(0, 40))
        kind = it % 4
        a = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
        if kind == 0:  # needle inside a longer haystack, lightly mutated
            pre = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 30))).astype(np.uint32)
            post = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 30))).astype(np.uint32)
            b = np.concatenate([pre, _gen.mutate(rng, a, int(rng.integers(0, 3)), alpha_n=alpha), post])
        elif kind == 1:  # equal lengths
            b = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
        else:  # independent lengths
            b = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 70))).astype(np.uint32)
        if it % 7 == 3:
            a, b = b, a
        out.append((a, b))
    return out


def _long_pairs(rng, lens, alpha=4):
    out = []
    for i, l1 in enumerate(lens):
        a = rng.integers(97, 97 + alpha, size=l1).astype(np.uint32)
        if i % 3 == 0:  # a mutated copy inside a longer haystack
            pre = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 200))).astype(np.uint32)
            post = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 200))).astype(np.uint32)
            b = np.concatenate([pre, _gen.mutate(rng, a, int(rng.integers(0, 1 + l1 // 20)), alpha_n=alpha), post])
        elif i % 3 == 1:  # equal lengths
            b = _gen.mutate(rng, a, int(rng.integers(0, 1 + l1 // 10)), alpha_n=alpha)[:l1]
            if len(b) < l1:
                b = np.concatenate([b, rng.integers(97, 97 + alpha, size=l1 - len(b)).astype(np.uint32)])
        else:
            b = rng.integers(97, 97 + alpha, size=l1 + int(rng.integers(1, 300))).astype(np.uint32)
        if i % 2:
            a, b = b, a
        out.append((a, b))
    return out


def _check(ctx, oracle, pairs, cutoff, width=None):
    import rapidfuzz_amd as rfa

    print(f"pairs {len(pairs)} cutoff {cutoff} width {width}", flush=True)
    a = ctx.strings([p[0] for p in pairs], width=width)
    b = ctx.strings([p[1] for p in pairs], width=width)
    got = rfa.partial_ratio_alignment(ctx, a, b, cutoff)
    a.close()
    b.close()
    for i, (s1, s2) in enumerate(pairs):
        exp = oracle.partial_ratio_alignment(s1, s2, cutoff)
        g = got[i]
        row = (float(g["score"]), int(g["src_start"]), int(g["src_end"]), int(g["dest_start"]), int(g["dest_end"]))
        assert np.float64(row[0]).view(np.uint64) == np.float64(exp[0]).view(np.uint64) and row[1:] == tuple(exp[1:]), \
            f"pair {i} ({len(s1)}, {len(s2)}) cutoff {cutoff}: got {row} expected {exp}"
    return got


def test_known_answer_issue_231(ctx, oracle):
    # tests-fuzz.cpp:187-200
    s1 = _u("er merkantilismus f/rderte handel und verkehr mit teils marktkonformen, teils "
            "dirigistischen ma_nahmen.")
    s2 = _u("ils marktkonformen, teils dirigistischen ma_nahmen. an der schwelle zum 19. "
            "jahrhundert entstand ein neu")
    got = _check(ctx, oracle, [(s1, s2), (s2, s1)], 0.0)
    assert abs(got[0]["score"] - 66.2337662) < 1e-6
    assert (int(got[0]["src_start"]), int(got[0]["src_end"]), int(got[0]["dest_start"]), int(got[0]["dest_end"])) == \
        (0, 103, 0, 51)


@pytest.mark.parametrize("cutoff", [0.0, 50.0, 75.0, 100.0, 100.5])
def test_small_random(ctx, oracle, cutoff):
    rng = np.random.default_rng(400 + int(cutoff))
    _check(ctx, oracle, _small_pairs(rng, 500), cutoff)


def test_cutoffs_at_reached_scores(ctx, oracle):
    """Each pair under its own reached score, the next double above and below it."""
    rng = np.random.default_rng(420)
    pairs = _small_pairs(rng, 150)
    base = _check(ctx, oracle, pairs, 0.0)
    for i, p in enumerate(pairs):
        s = float(base[i]["score"])
        for c in (s, float(np.nextafter(s, 200.0)), float(np.nextafter(s, 0.0))):
            _check(ctx, oracle, [p], c)


@pytest.mark.parametrize("width", [1, 2, 4])
def test_unit_widths(ctx, oracle, width):
    rng = np.random.default_rng(430 + width)
    pairs = _small_pairs(rng, 200)
    if width == 2:
        pairs = [(a + 60000 - 97, b + 60000 - 97) for a, b in pairs]
    elif width == 4:
        pairs = [(a + (1 << 24), b + (1 << 24)) for a, b in pairs]
    _check(ctx, oracle, pairs, 0.0, width=width)


@pytest.mark.parametrize("lens", [(1, 2, 63, 64, 65), (100, 128, 200, 255, 256), (257, 400, 640, 1000, 1024)])
@pytest.mark.parametrize("cutoff", [0.0, 80.0])
def test_long_needles(ctx, oracle, lens, cutoff):
    rng = np.random.default_rng(440 + lens[0] + int(cutoff))
    _check(ctx, oracle, _long_pairs(rng, list(lens) * 2), cutoff)


@pytest.mark.parametrize("l1", [5, 40, 64, 100, 200, 300])
def test_many_equal_windows(ctx, oracle, l1):
    """Haystacks of runs and repeats, where many full windows share the smallest distance and
    the search's visiting order decides the reported window."""
    rng = np.random.default_rng(460 + l1)
    pairs = []
    for i in range(24):
        nd = rng.integers(97, 99, size=l1).astype(np.uint32)
        if i % 3 == 0:
            hy = np.full(l1 + int(rng.integers(1, 400)), 97, np.uint32)
        elif i % 3 == 1:
            hy = np.tile(nd, 2 + i % 5)[: l1 + int(rng.integers(1, 3 * l1 + 2))]
        else:
            hy = np.concatenate([np.full(int(rng.integers(0, 150)), 98, np.uint32), nd,
                                 np.full(int(rng.integers(1, 150)), 97, np.uint32)])
        pairs.append((nd, hy) if i % 2 else (hy, nd))
    for c in (0.0, 60.0):
        _check(ctx, oracle, pairs, c)


def test_trivial_pairs(ctx, oracle):
    e = np.zeros(0, np.uint32)
    pairs = [(e, e), (e, _u("abc")), (_u("abc"), e), (_u("abc"), _u("abc")), (_u("a"), _u("b")),
             (_u("ab"), _u("xaby")), (_u("xaby"), _u("ab"))]
    for c in (0.0, 100.0, 100.5, -5.0):
        _check(ctx, oracle, pairs, c)


@pytest.mark.parametrize("cutoff", [0.0, 85.0])
def test_needles_past_1024(ctx, oracle, cutoff):
    """Shorter strings of 1,025-4,096 units (one wave per pair, the reference's bisection window
    by window), in a batch with short pairs; both orders, equal lengths, 8/16-bit units."""
    rng = np.random.default_rng(480 + int(cutoff))
    pairs = _long_pairs(rng, [1025, 1300, 2048, 3000, 4096, 1100], alpha=4)
    pairs += _small_pairs(rng, 40)
    _check(ctx, oracle, pairs, cutoff)
    wide = [(x + 0x4E00 - 97, y + 0x4E00 - 97) for x, y in pairs[:6]]
    _check(ctx, oracle, wide, cutoff, width=2)


def test_length_limit(ctx):
    import rapidfuzz_amd as rfa

    a = ctx.strings([np.full(4097, 97, np.uint32)])
    b = ctx.strings([np.full(4100, 97, np.uint32)])
    with pytest.raises(rfa.RfaError):
        rfa.partial_ratio_alignment(ctx, a, b, 0.0)
    # a longer haystack is fine while the shorter string fits
    c = ctx.strings([np.full(4096, 97, np.uint32)])
    got = rfa.partial_ratio_alignment(ctx, c, a, 0.0)
    assert float(got[0]["score"]) == 100.0
