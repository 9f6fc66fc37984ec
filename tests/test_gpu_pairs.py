"""GPU parity of rfa_pairs (independent pairs, mixed metrics) against the oracle."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu


def _pairs(rng, n, lmax, alpha_n, width):
    a, b = [], []
    for _ in range(n):
        la = int(min(lmax, max(0, rng.lognormal(np.log(40), 1.0))))
        s = rng.integers(97, 97 + alpha_n, size=la).astype(np.uint32)
        if width == 4:
            s = (s * 40503 + 7).astype(np.uint32)
        if rng.random() < 0.5:
            t = _gen.mutate(rng, s, int(rng.integers(0, 8)), alpha_lo=97, alpha_n=alpha_n)
            if width == 4:
                t = np.where(t < 200, t * 40503 + 7, t).astype(np.uint32)
        else:
            lb = int(min(lmax, max(0, rng.lognormal(np.log(40), 1.0))))
            t = rng.integers(97, 97 + alpha_n, size=lb).astype(np.uint32)
            if width == 4:
                t = (t * 40503 + 7).astype(np.uint32)
        a.append(s)
        b.append(t)
    return a, b


@pytest.mark.parametrize("schedule", [0, 1])
@pytest.mark.parametrize("width", [1, 4])
@pytest.mark.parametrize("kind,cu,cd", [(0, None, None), (0, 5, None), (1, 10, None), (2, None, 0.3),
                                        (3, None, 0.7)])
def test_pairs(ctx, oracle, schedule, width, kind, cu, cd):
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(11 + width + 7 * schedule + 13 * kind)
    a, b = _pairs(rng, 300, 600, 5, width)
    metrics = rng.integers(0, 4, size=len(a)).astype(np.int32)
    A = ctx.strings(a, width=width)
    B = ctx.strings(b, width=width)
    if cu is None:
        cu = rfa.INT64_MAX if kind == 0 else 0
    if cd is None:
        cd = 1.0 if kind == 2 else 0.0
    got = rfa.pairs(ctx, A, B, metrics, kind=kind, cutoff_u=cu, cutoff_d=cd, schedule=schedule)
    exp = oracle.pairs(metrics, a, b, kind=kind, cutoff_u=cu, cutoff_d=cd)
    if got.dtype == np.float64:
        bad = np.flatnonzero(got.view(np.uint64) != exp.view(np.uint64))
    else:
        bad = np.flatnonzero(got != exp)
    assert bad.size == 0, (f"{bad.size} mismatches; first {bad[0]}: metric {metrics[bad[0]]} "
                           f"lens {len(a[bad[0]])}/{len(b[bad[0]])} got {got[bad[0]]} expected {exp[bad[0]]}")


def test_pairs_schedules_agree_on_long_tail(ctx):
    """Both schedules on pairs up to 2000 units (the long-pattern block path)."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(5)
    a, b = _pairs(rng, 200, 2000, 4, 1)
    a += [np.full(1500, 97, np.uint32), np.arange(2000, dtype=np.uint32) % 4 + 97]
    b += [np.full(1400, 98, np.uint32), (np.arange(1990, dtype=np.uint32) * 7) % 4 + 97]
    metrics = rng.integers(0, 4, size=len(a)).astype(np.int32)
    A, B = ctx.strings(a, width=1), ctx.strings(b, width=1)
    x = rfa.pairs(ctx, A, B, metrics, schedule=0)
    y = rfa.pairs(ctx, A, B, metrics, schedule=1)
    assert np.array_equal(x, y)


def test_pairs_rejects_bad_metric(ctx):
    import rapidfuzz_amd as rfa

    A = ctx.strings([b"ab"])
    with pytest.raises(rfa.InvalidArgument):
        rfa.pairs(ctx, A, A, [5])


def test_pairs_rejects_count_mismatch(ctx):
    import rapidfuzz_amd as rfa

    with pytest.raises(rfa.InvalidArgument):
        rfa.pairs(ctx, ctx.strings([b"ab", b"c"]), ctx.strings([b"ab"]), [0, 0])


def _class_pairs(rng, width, sym_hi):
    """Patterns across every lanes-per-pair class (1..64 words) and the block boundary
    (4096 / 4097 units), texts up to twice as long, symbols below sym_hi."""
    lens = [0, 1, 5, 31, 32, 33, 63, 64, 65, 127, 128, 129, 200, 255, 256, 257, 511, 512, 513, 1000, 1023, 1024,
            1025, 2047, 2048, 2049, 3000, 4095, 4096, 4097]
    a, b = [], []
    lo = max(0, sym_hi - 6)  # a small alphabet near the top of the range, so every symbol bit matters
    for m in lens:
        for _ in range(3):
            s = rng.integers(lo, sym_hi, size=m).astype(np.uint64)
            if rng.random() < 0.5:
                t = _gen.mutate(rng, s - lo, int(rng.integers(0, 12)), alpha_lo=0, alpha_n=sym_hi - lo).astype(np.uint64) + lo
            else:
                t = rng.integers(lo, sym_hi, size=int(rng.integers(m, 2 * m + 3))).astype(np.uint64)
            # a few symbols of the wider range on both sides
            if m and rng.random() < 0.5:
                s[rng.integers(0, m)] = rng.integers(0, sym_hi)
            if rng.random() < 0.5:
                a.append(s)
                b.append(t)
            else:
                a.append(t)
                b.append(s)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[width]
    return [x.astype(dt).astype(np.uint32) for x in a], [x.astype(dt).astype(np.uint32) for x in b]


@pytest.mark.parametrize("width,sym_hi", [(1, 256), (1, 100), (2, 200), (2, 65536), (4, 250), (4, 65536),
                                          (4, 1 << 20), (4, 1 << 31)])
@pytest.mark.parametrize("kind,cu,cd", [(0, None, None), (0, 40, None), (1, 30, None), (2, None, 0.4),
                                        (3, None, 0.6)])
def test_pairs_segment_classes(ctx, oracle, width, sym_hi, kind, cu, cd):
    """The segmented kernel's classes (pattern words 1..64), symbol bit counts 8/16/24/32,
    metrics mixed inside a class, OSA and > 4096-unit patterns routed elsewhere."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(sym_hi % 1000 + width + 17 * kind)
    a, b = _class_pairs(rng, width, sym_hi)
    metrics = rng.integers(0, 3, size=len(a)).astype(np.int32)
    metrics[rng.random(len(a)) < 0.1] = 3
    A = ctx.strings(a, width=width)
    B = ctx.strings(b, width=width)
    if cu is None:
        cu = rfa.INT64_MAX if kind == 0 else 0
    if cd is None:
        cd = 1.0 if kind == 2 else 0.0
    got = rfa.pairs(ctx, A, B, metrics, kind=kind, cutoff_u=cu, cutoff_d=cd)
    exp = oracle.pairs(metrics, a, b, kind=kind, cutoff_u=cu, cutoff_d=cd)
    if got.dtype == np.float64:
        bad = np.flatnonzero(got.view(np.uint64) != exp.view(np.uint64))
    else:
        bad = np.flatnonzero(got != exp)
    assert bad.size == 0, (f"{bad.size} mismatches; first {bad[0]}: metric {metrics[bad[0]]} "
                           f"lens {len(a[bad[0]])}/{len(b[bad[0]])} got {got[bad[0]]} expected {exp[bad[0]]}")


def test_pairs_many_short_mixed(ctx, oracle):
    """Thousands of short pairs: full 64-pair waves, metric boundaries inside waves, ragged tails."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(99)
    a, b = _pairs(rng, 5000, 300, 6, 4)
    metrics = rng.integers(0, 3, size=len(a)).astype(np.int32)
    A, B = ctx.strings(a, width=4), ctx.strings(b, width=4)
    got = rfa.pairs(ctx, A, B, metrics)
    exp = oracle.pairs(metrics, a, b)
    assert np.array_equal(got, exp)
    plain = rfa.pairs(ctx, A, B, metrics, schedule=rfa.SCHEDULE_PLAIN)
    assert np.array_equal(plain, exp)


@pytest.mark.parametrize("with_block", [False, True])
def test_pairs_graph_replays(ctx, oracle, with_block):
    """Repeated identical rfa_pairs calls are captured once into a HIP graph and replayed.  Every
    replay must equal the oracle, with eager work (other pairs calls, a cdist, synchronizations)
    between the replays, new metric values between replays, and a larger call in between that
    regrows the plan buffers (which forces a new capture)."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(123 + with_block)
    a, b = _pairs(rng, 700, 600, 5, 4)
    if with_block:  # block kernel work (> 4096-unit patterns, long OSA) overlapping on the second stream
        a += [rng.integers(97, 102, size=4200).astype(np.uint32), rng.integers(97, 102, size=300).astype(np.uint32)]
        b += [rng.integers(97, 102, size=4300).astype(np.uint32), rng.integers(97, 102, size=280).astype(np.uint32)]
    n = len(a)
    A, B = ctx.strings(a, width=4), ctx.strings(b, width=4)
    other_a, other_b = _pairs(rng, 300, 200, 5, 4)
    OA, OB = ctx.strings(other_a, width=4), ctx.strings(other_b, width=4)
    big_a, big_b = _pairs(rng, 3000, 300, 5, 4)
    BA, BB = ctx.strings(big_a, width=4), ctx.strings(big_b, width=4)
    sc = rfa.Scorer(ctx, 0, ctx.strings(a[:20], width=4))
    ctx.set_option(rfa.OPTION_PAIRS_GRAPH, 1)
    try:
        _graph_replays(ctx, oracle, rng, A, B, a, b, OA, OB, other_a, other_b, BA, BB, big_a, big_b, sc, with_block)
    finally:
        ctx.set_option(rfa.OPTION_PAIRS_GRAPH, 0)


def _graph_replays(ctx, oracle, rng, A, B, a, b, OA, OB, other_a, other_b, BA, BB, big_a, big_b, sc, with_block):
    import rapidfuzz_amd as rfa

    n = len(a)
    for it in range(8):
        metrics = rng.integers(0, 3, size=n).astype(np.int32)
        if with_block:
            metrics[-1] = 3  # the long OSA pair
        if it < 3:
            metrics = np.resize(np.array([0, 1, 2], np.int32), n)
            if with_block:
                metrics[-1] = 3
        got = rfa.pairs(ctx, A, B, metrics, kind=0, cutoff_u=rfa.INT64_MAX)
        exp = oracle.pairs(metrics, a, b)
        assert np.array_equal(got, exp), f"replay {it}: {np.count_nonzero(got != exp)} differ"
        # eager work between the replays
        om = rng.integers(0, 3, size=len(other_a)).astype(np.int32)
        assert np.array_equal(rfa.pairs(ctx, OA, OB, om), oracle.pairs(om, other_a, other_b))
        sc.cdist(OB, 0)
        ctx.synchronize()
        if it == 5:
            bm = rng.integers(0, 3, size=len(big_a)).astype(np.int32)
            assert np.array_equal(rfa.pairs(ctx, BA, BB, bm), oracle.pairs(bm, big_a, big_b))


def test_pairs_chain_threshold_and_autotune(ctx, oracle):
    """Every long-pair threshold (which string is the pattern, which pairs run first) gives the
    same results; the autotuner picks one of its candidates and keeps it in the context."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(321)
    a, b = _pairs(rng, 600, 3000, 5, 4)
    a += [rng.integers(97, 102, size=3500).astype(np.uint32), rng.integers(97, 102, size=20).astype(np.uint32)]
    b += [rng.integers(97, 102, size=40).astype(np.uint32), rng.integers(97, 102, size=4000).astype(np.uint32)]
    metrics = rng.integers(0, 3, size=len(a)).astype(np.int32)
    A, B = ctx.strings(a, width=4), ctx.strings(b, width=4)
    exp = oracle.pairs(metrics, a, b)
    try:
        for chain in (1, 64, 128, 512, 2048, 4097, 1 << 31):
            ctx.set_option(rfa.OPTION_PAIRS_CHAIN, chain)
            got = rfa.pairs(ctx, A, B, metrics)
            assert np.array_equal(got, exp), f"chain {chain}: {np.count_nonzero(got != exp)} differ"
        chosen = rfa.pairs_autotune(ctx, A, B, metrics)
        assert chosen in (128, 256, 512, 1024, 2048, 4097)
        assert np.array_equal(rfa.pairs(ctx, A, B, metrics), exp)
    finally:
        ctx.set_option(rfa.OPTION_PAIRS_CHAIN, 512)
    with pytest.raises(rfa.InvalidArgument):
        ctx.set_option(rfa.OPTION_PAIRS_CHAIN, 0)
