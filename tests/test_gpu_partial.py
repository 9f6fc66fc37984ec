"""GPU parity of fuzz::partial_ratio (rfa_cdist with RFA_PARTIAL_RATIO) against the
oracle, bit for bit, and against the reference's own assertions."""
import numpy as np
import pytest

import _gen
import golden_cases as G

pytestmark = pytest.mark.gpu

PARTIAL, SIM = 10, 1
CUTS = (0.0, 40.0, 70.0, 90.0, 100.0, 100.5)


def _check(ctx, oracle, queries, choices, width=None, cuts=CUTS):
    import rapidfuzz_amd as rfa

    sc = rfa.Scorer(ctx, PARTIAL, ctx.strings(queries, width=width))
    cs = ctx.strings(choices, width=width)
    for cut in cuts:
        got = sc.cdist(cs, SIM, cutoff_d=cut)
        exp = oracle.cdist(PARTIAL, SIM, queries, choices, cutoff_d=cut)
        bad = np.argwhere(got.view(np.uint64) != exp.view(np.uint64))
        if len(bad):
            i, j = bad[0]
            raise AssertionError(f"cutoff {cut}: {len(bad)} differ; q{i} (len {len(queries[i])}) c{j} "
                                 f"(len {len(choices[j])}): got {got[i, j]!r} expected {exp[i, j]!r}")


def test_known_answers(ctx):
    """tests-fuzz.cpp assertions (within their 1e-6), both argument orders, including the
    103/104-, 65/65- and 3009/364-unit cases of issues 231, 257 and 219."""
    import rapidfuzz_amd as rfa

    for c in G.load()["cases"]:
        if c["metric"] != PARTIAL:
            continue
        a, b = G.expand(c["s1"]), G.expand(c["s2"])
        sc = rfa.Scorer(ctx, PARTIAL, ctx.strings([a]))
        got = float(sc.cdist(ctx.strings([b]), SIM, cutoff_d=c.get("cutoff", 0.0))[0, 0])
        assert abs(got - c["expected"]) <= c["tol"], (G.case_id(c), got, c["expected"])


@pytest.mark.parametrize("width", [1, 2, 4])
def test_short_queries(ctx, oracle, width):
    """Tile path: queries of 0-64 units (32- and 64-bit tiles) against choices of 0-130
    units, needles on both sides, equal lengths, embedded near-copies."""
    rng = np.random.default_rng(900 + width)
    base = {1: 97, 2: 40000, 4: 3000000}[width]
    q = [x + (base - 97) for x in _gen.random_strings(rng, 90, 0, 64, alpha_n=5)]
    q += [x + (base - 97) for x in _gen.random_strings(rng, 40, 0, 30, alpha_n=3)]
    c = [x + (base - 97) for x in _gen.random_strings(rng, 50, 0, 130, alpha_n=5)]
    for i in range(30):
        s = q[i]
        pre = rng.integers(base, base + 5, size=int(rng.integers(0, 40))).astype(np.uint32)
        post = rng.integers(base, base + 5, size=int(rng.integers(0, 40))).astype(np.uint32)
        c.append(np.concatenate([pre, _gen.mutate(rng, s - (base - 97), 2) + (base - 97), post]))
        c.append(s.copy())                                  # equal length, equal
        c.append(_gen.mutate(rng, s - (base - 97), 1)[:len(s)] + (base - 97))  # about equal length
    _check(ctx, oracle, q, c, width=width, cuts=(0.0, 50.0, 85.0, 100.5))


@pytest.mark.parametrize("width", [1, 2, 4])
def test_long_queries(ctx, oracle, width):
    """Long path: queries of 65-300 units against choices of 0-400 units (the choice as
    the needle, the query as the needle, equal lengths)."""
    rng = np.random.default_rng(950 + width)
    base = {1: 97, 2: 50000, 4: 1000000}[width]
    q = [x + (base - 97) for x in _gen.random_strings(rng, 5, 65, 300, alpha_n=4)]
    c = [x + (base - 97) for x in _gen.random_strings(rng, 10, 0, 400, alpha_n=4)]
    c += [q[0][10:80].copy(), q[1].copy(), q[2][:64].copy(), q[3][:65].copy(), q[4][:1].copy()]
    c.append(np.concatenate([c[0][:50], _gen.mutate(rng, q[0] - (base - 97), 5) + (base - 97)]))
    _check(ctx, oracle, q, c, width=width, cuts=(0.0, 60.0, 95.0))


@pytest.mark.parametrize("width", [1, 4])
def test_needles_past_1024(ctx, oracle, width):
    """Pairs whose shorter string has 1,025-3,000 units (one wave per pair) beside pairs the
    long kernel takes, in one cdist: queries of 1,100-3,000 units against choices of 0-3,500."""
    rng = np.random.default_rng(960 + width)
    base = {1: 97, 4: 1000000}[width]
    q = [x + (base - 97) for x in _gen.random_strings(rng, 3, 1100, 3000, alpha_n=4)]
    q.append(_gen.random_strings(rng, 1, 100, 200, alpha_n=4)[0] + (base - 97))
    c = [x + (base - 97) for x in _gen.random_strings(rng, 6, 0, 3500, alpha_n=4)]
    c += [np.concatenate([c[0][:300], _gen.mutate(rng, q[0] - (base - 97), 20) + (base - 97), c[1][:100]]),
          q[1][50:1500].copy(), q[2].copy()]
    _check(ctx, oracle, q, c, width=width, cuts=(0.0, 80.0))


def test_large_alphabet_tiles(ctx, oracle):
    """Short queries whose tile has more symbols than the LDS table holds take the long path."""
    rng = np.random.default_rng(990)
    q = [rng.integers(0, 60000, size=int(rng.integers(1, 40))).astype(np.uint32) for _ in range(70)]
    c = [rng.integers(0, 60000, size=int(rng.integers(0, 60))).astype(np.uint32) for _ in range(20)]
    c += [np.concatenate([c[i][:5], q[i], c[i][5:9]]) for i in range(10)]
    _check(ctx, oracle, q, c, width=2, cuts=(0.0, 80.0))


def test_rejected_calls(ctx):
    import rapidfuzz_amd as rfa

    qs = ctx.strings([b"abc"])
    sc = rfa.Scorer(ctx, PARTIAL, qs)
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"abcd"]), 0)  # partial_ratio only has similarity
    with pytest.raises(rfa.InvalidArgument):
        sc.cdist(ctx.strings([b"abcd"]), SIM, semantics=1)  # no Multi* scorer
    long_q = rfa.Scorer(ctx, PARTIAL, ctx.strings([b"a" * 4100]))
    with pytest.raises(rfa.RfaError):
        long_q.cdist(ctx.strings([b"b" * 4100]), SIM)  # both strings beyond the 4096-unit needle limit


def test_topk_partial(ctx, oracle):
    """extract()-style top-k over partial_ratio (the dense path), (score desc, index asc)."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(995)
    q = _gen.random_strings(rng, 6, 3, 20, alpha_n=4)
    c = _gen.random_strings(rng, 300, 0, 80, alpha_n=4)
    sc = rfa.Scorer(ctx, PARTIAL, ctx.strings(q))
    idx, score = sc.topk(ctx.strings(c), 5, cutoff_d=60.0)
    full = oracle.cdist(PARTIAL, SIM, q, c, cutoff_d=60.0)
    for i in range(len(q)):
        order = sorted(((-full[i, j], j) for j in range(len(c)) if full[i, j] >= 60.0))[:5]
        exp_idx = [j for _, j in order] + [-1] * (5 - len(order))
        exp_sc = [-s for s, _ in order] + [0.0] * (5 - len(order))
        assert list(idx[i]) == exp_idx
        assert np.array_equal(score[i], np.array(exp_sc))
