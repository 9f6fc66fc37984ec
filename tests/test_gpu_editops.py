"""GPU parity of rfa_editops (levenshtein_editops, lcs_seq_editops / indel_editops)
against the oracle, operation for operation: random pairs around the 64-unit single
word of the stripped first string (Levenshtein past it up to 4096 units through the
block band and the Hirschberg split, LCS up to 4096 units over several words),
long second strings (several 64-unit chunks of rows),
common affixes, empty strings, 8/16/32-bit units, and the reference's editops test
strings; every script also rebuilds the second string."""
import numpy as np
import pytest

import _gen
from test_editops_oracle import REF_STRINGS, _codes, apply_ops

pytestmark = pytest.mark.gpu

LEV, INDEL, LCS, HAMMING = 0, 1, 2, 7


def _pairs(rng, alpha, n=120):
    out = [(_codes(a), _codes(b)) for a, b in REF_STRINGS]
    out += [(np.zeros(0, np.uint32), np.zeros(0, np.uint32)), (np.zeros(0, np.uint32), _codes("abc")),
            (_codes("abc"), np.zeros(0, np.uint32)), (_codes("abc"), _codes("abc"))]
    for _ in range(n):
        s1 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 65))).astype(np.uint32)
        s2 = rng.integers(97, 97 + alpha, size=int(rng.integers(0, 300))).astype(np.uint32)
        out.append((s1, s2))
        out.append((s1, _gen.mutate(rng, s1, int(rng.integers(0, 8)), alpha_n=alpha)))
    # a common affix around a 64-unit core
    core = rng.integers(97, 97 + alpha, size=64).astype(np.uint32)
    pre, suf = _codes("prefix-"), _codes("-suffix")
    out.append((np.concatenate([pre, core, suf]), np.concatenate([pre, _gen.mutate(rng, core, 5), suf])))
    return out


def check(ctx, oracle, metric, pairs, width=1):
    import rapidfuzz_amd as rfa

    print(f"metric {metric} width {width} pairs {len(pairs)}", flush=True)
    a = ctx.strings([p[0] for p in pairs], width=width)
    b = ctx.strings([p[1] for p in pairs], width=width)
    off, ops = rfa.editops(ctx, a, b, metric)
    for i, (s1, s2) in enumerate(pairs):
        got = ops[int(off[i]):int(off[i + 1])]
        exp = oracle.editops(metric, s1, s2)
        assert np.array_equal(got["type"], exp[0]) and np.array_equal(got["src_pos"], exp[1]) and \
            np.array_equal(got["dest_pos"], exp[2]), f"pair {i} ({len(s1)}, {len(s2)}): {got} vs {exp}"
        assert apply_ops((got["type"], got["src_pos"], got["dest_pos"]), s1, s2) == [int(x) for x in s2]
    a.close()
    b.close()


@pytest.mark.parametrize("metric", [LEV, INDEL, LCS, HAMMING])
@pytest.mark.parametrize("alpha", [2, 5])
def test_random(ctx, oracle, metric, alpha):
    check(ctx, oracle, metric, _pairs(np.random.default_rng(10 * metric + alpha), alpha))


@pytest.mark.parametrize("width", [2, 4])
@pytest.mark.parametrize("metric", [LEV, LCS, HAMMING])
def test_wide_units(ctx, oracle, width, metric):
    rng = np.random.default_rng(30 + width)
    lo = 0x4E00 if width == 2 else 0x1F300
    pairs = [(rng.integers(lo, lo + 6, size=int(rng.integers(0, 65))).astype(np.uint32),
              rng.integers(lo, lo + 6, size=int(rng.integers(0, 130))).astype(np.uint32)) for _ in range(80)]
    check(ctx, oracle, metric, pairs, width=width)


@pytest.mark.parametrize("metric", [INDEL, LCS])
def test_lcs_long_first_strings(ctx, oracle, metric):
    """LCS editops over several words: stripped first strings of 65-4096 units."""
    rng = np.random.default_rng(50 + metric)
    pairs = []
    for n1 in (65, 100, 128, 129, 300, 700, 1500, 4096):
        s1 = rng.integers(97, 101, size=n1).astype(np.uint32)
        pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(1, 30)))))
        pairs.append((s1, rng.integers(97, 101, size=int(rng.integers(0, 400))).astype(np.uint32)))
    pairs += [(rng.integers(97, 99, size=int(rng.integers(65, 200))).astype(np.uint32),
               rng.integers(97, 99, size=int(rng.integers(0, 200))).astype(np.uint32)) for _ in range(30)]
    pairs += _pairs(rng, 3, n=20)
    check(ctx, oracle, metric, pairs)


@pytest.mark.parametrize("width", [1, 2])
def test_levenshtein_long_first_strings(ctx, oracle, width):
    """Levenshtein editops past 64 stripped units: the reference's block band (first strings
    of 65-4096 units below its Hirschberg split, second strings under 10 units at any length)."""
    from test_editops_oracle import _long_lev_pairs

    rng = np.random.default_rng(70 + width)
    pairs = _long_lev_pairs(rng, 120)
    # word boundaries, common affixes around a long core, both orders
    for n1 in (65, 127, 128, 129, 191, 192, 1023, 1024, 1025):
        s1 = rng.integers(97, 100, size=n1).astype(np.uint32)
        pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(1, 12)), alpha_n=3)))
        pairs.append((_gen.mutate(rng, s1, 3, alpha_n=3), s1))
    core = rng.integers(97, 100, size=300).astype(np.uint32)
    pairs.append((np.concatenate([_codes("prefix-"), core, _codes("-suffix")]),
                  np.concatenate([_codes("prefix-"), _gen.mutate(rng, core, 20, alpha_n=3), _codes("-suffix")])))
    pairs += _pairs(rng, 4, n=30)  # short pairs in the same batch
    if width == 2:
        pairs = [(x + 0x4E00 - 97, y + 0x4E00 - 97) for x, y in pairs]
    check(ctx, oracle, LEV, pairs, width=width)


@pytest.mark.parametrize("width", [1, 2])
def test_levenshtein_hirschberg(ctx, oracle, width):
    """Pairs the reference splits by Hirschberg (stripped |s1| > 64, |s2| >= 10 and
    |s1| * |s2| >= 2^22; Levenshtein_impl.hpp:1185-1212): mutated copies, independent
    strings, shifted copies, short first strings against long second ones (leaves under
    65 units), first strings up to the 4096-unit limit against longer second strings, and
    short pairs in the same batch."""
    rng = np.random.default_rng(90 + width)
    pairs = []
    for it in range(12):
        alpha = int(rng.choice([2, 4, 26]))
        n1 = int(rng.integers(2050, 4000))
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
    s1 = rng.integers(97, 101, size=100).astype(np.uint32)
    pairs.append((s1, rng.integers(97, 101, size=45000).astype(np.uint32)))
    s1 = rng.integers(97, 101, size=4096).astype(np.uint32)
    pairs.append((s1, np.concatenate([_gen.mutate(rng, s1, 200, alpha_n=4),
                                      rng.integers(97, 101, size=4000).astype(np.uint32)])))
    core = rng.integers(97, 99, size=3000).astype(np.uint32)
    pairs.append((np.concatenate([_codes("prefix-"), core, _codes("-suffix")]),
                  np.concatenate([_codes("prefix-"), _gen.mutate(rng, core, 300, alpha_n=2), _codes("-suffix")])))
    pairs += _pairs(rng, 4, n=20)
    if width == 2:
        pairs = [(x + 0x4E00 - 97, y + 0x4E00 - 97) for x, y in pairs]
    check(ctx, oracle, LEV, pairs, width=width)


def test_hamming_long(ctx, oracle):
    """Hamming editops have no length limit: strings of up to 5000 units."""
    rng = np.random.default_rng(61)
    pairs = [(rng.integers(97, 99, size=n1).astype(np.uint32), rng.integers(97, 99, size=n2).astype(np.uint32))
             for n1, n2 in ((5000, 5000), (4000, 100), (70, 3000), (0, 129), (129, 0))]
    check(ctx, oracle, HAMMING, pairs)


def test_sizes_then_ops_and_limits(ctx, oracle):
    import rapidfuzz_amd as rfa

    a = ctx.strings([b"Lorem ipsum.", b"aa"])
    b = ctx.strings([b"XYZLorem ABC iPsum", b"abb"])
    off = np.zeros(3, np.uint64)
    assert rfa._lib.rfa_editops(ctx.h, a.h, b.h, 0, off.ctypes.data, None, 0) == 0
    small = np.zeros(1, dtype=rfa.EDITOP_DTYPE)
    assert rfa._lib.rfa_editops(ctx.h, a.h, b.h, 0, off.ctypes.data, small.ctypes.data, 1) == 1  # too small
    assert int(off[2]) == len(oracle.editops(0, b"Lorem ipsum.", b"XYZLorem ABC iPsum")[0]) + \
        len(oracle.editops(0, b"aa", b"abb")[0])
    # Levenshtein pairs the reference splits (Hirschberg) are supported
    check(ctx, oracle, LEV, [(np.full(2100, 97, np.uint32), np.full(2100, 98, np.uint32))])
    over_a = ctx.strings([np.full(4097, 97, np.uint32)], width=1)
    over_b = ctx.strings([np.full(5, 98, np.uint32)], width=1)
    with pytest.raises(rfa.RfaError):
        rfa.editops(ctx, over_a, over_b, 0)  # Levenshtein: more than 4096 units after the affixes
    huge_a = ctx.strings([np.full(4097, 97, np.uint32)], width=1)
    huge_b = ctx.strings([np.full(10, 98, np.uint32)], width=1)
    with pytest.raises(rfa.RfaError):
        rfa.editops(ctx, huge_a, huge_b, 2)  # LCS: more than 4096 units after the affixes
    with pytest.raises(rfa.InvalidArgument):
        rfa.editops(ctx, a, b, 3)  # OSA has no editops
