"""rfa_pairs_noalloc (the entry point that allocates nothing) against rfa_pairs, bit for bit:
two calls in a row on one context, the second with fewer pairs, with and without OSA codes,
integer and normalized kinds."""
import numpy as np
import pytest

import _gen

pytestmark = pytest.mark.gpu


def _pairs(rng, n, lmax, alpha_n):
    a, b = [], []
    for _ in range(n):
        la = int(min(lmax, max(0, rng.lognormal(np.log(30), 0.9))))
        s = rng.integers(97, 97 + alpha_n, size=la).astype(np.uint32)
        if rng.random() < 0.5:
            t = _gen.mutate(rng, s, int(rng.integers(0, 8)), alpha_lo=97, alpha_n=alpha_n)
        else:
            t = rng.integers(97, 97 + alpha_n, size=int(min(lmax, max(0, rng.lognormal(np.log(30), 0.9))))).astype(np.uint32)
        a.append(s)
        b.append(t)
    return a, b


@pytest.mark.parametrize("has_osa", [0, 1])
@pytest.mark.parametrize("kind,cu,cd", [(0, None, None), (1, 4, None), (3, None, 0.7)])
def test_noalloc_matches_pairs(ctx, has_osa, kind, cu, cd):
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(17 + has_osa + 3 * kind)
    lmax = 64 if has_osa else 3000
    cu = rfa.INT64_MAX if cu is None and kind == 0 else (0 if cu is None else cu)
    cd = 1.0 if cd is None else cd
    for n in (20000, 7000):
        a, b = _pairs(rng, n, lmax, 6)
        metrics = rng.integers(0, 4 if has_osa else 3, size=n).astype(np.int32)
        A, B = ctx.strings(a, width=1), ctx.strings(b, width=1)
        exp = rfa.pairs(ctx, A, B, metrics, kind=kind, cutoff_u=cu, cutoff_d=cd)
        dm = hip.DeviceBuffer(n)
        dm.from_host(metrics.astype(np.uint8))
        dout = hip.DeviceBuffer(n * 8)
        rfa.pairs_noalloc(ctx, A, B, dm.ptr.value, has_osa, dout.ptr.value, kind=kind, cutoff_u=cu, cutoff_d=cd)
        ctx.synchronize()
        got = np.zeros(n, dtype=exp.dtype)
        dout.to_host(got)
        g = got.view(np.uint64)
        e = exp.view(np.uint64)
        bad = np.flatnonzero(g != e)
        assert bad.size == 0, f"n={n}: {bad.size} differ; first {bad[0]}: got {got[bad[0]]} expected {exp[bad[0]]}"
        dm.free()
        dout.free()
        A.close()
        B.close()


def test_noalloc_rejects_beyond_limits(ctx):
    """Refused on the host, before any launch."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    dm, dout = hip.DeviceBuffer(64), hip.DeviceBuffer(64)
    dm.from_host(np.zeros(64, np.uint8))
    A = ctx.strings([np.zeros(5000, np.uint32)], width=1)
    with pytest.raises(rfa.RfaError):
        rfa.pairs_noalloc(ctx, A, A, dm.ptr.value, 0, dout.ptr.value)  # shorter string over 4096 units
    C = ctx.strings([np.zeros(100, np.uint32)], width=1)
    with pytest.raises(rfa.RfaError):
        rfa.pairs_noalloc(ctx, C, C, dm.ptr.value, 1, dout.ptr.value)  # OSA codes with a 100-unit shorter string
    many = ctx.strings([np.zeros(1, np.uint32)] * ((1 << 20) + 1), width=1)
    with pytest.raises(rfa.RfaError):
        rfa.pairs_noalloc(ctx, many, many, dm.ptr.value, 0, dout.ptr.value)  # more than 2^20 pairs
    dm.free()
    dout.free()
