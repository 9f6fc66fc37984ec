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
