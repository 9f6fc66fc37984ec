"""The HIP path against the known answers of the reference's own unit tests."""
import numpy as np
import pytest

import golden_cases as G

pytestmark = pytest.mark.gpu
DATA = G.load()


def _score(ctx, metric, kind, s1, s2, weights=(1, 1, 1), cutoff=None, semantics=0):
    import rapidfuzz_amd as rfa

    width = 1 if max([int(s1.max()) if s1.size else 0, int(s2.max()) if s2.size else 0]) < 256 else 4
    sc = rfa.Scorer(ctx, metric, ctx.strings([s1], width=width), weights=weights)
    kw = {}
    if cutoff is not None:
        if rfa.is_int_result(metric, kind):
            kw["cutoff_u"] = cutoff
        else:
            kw["cutoff_d"] = cutoff
    return sc.cdist(ctx.strings([s2], width=width), kind, semantics=semantics, **kw)[0, 0]


@pytest.mark.parametrize("c", DATA["cases"], ids=[G.case_id(c) for c in DATA["cases"]])
def test_case(ctx, c):
    got = _score(ctx, c["metric"], c["kind"], G.expand(c["s1"]), G.expand(c["s2"]),
                 weights=tuple(c.get("weights", (1, 1, 1))), cutoff=c.get("cutoff"))
    if "tol" in c:
        assert abs(float(got) - c["expected"]) <= c["tol"]
    else:
        assert got == c["expected"]


@pytest.mark.parametrize("c", DATA["multi"], ids=[G.case_id(c) for c in DATA["multi"]])
def test_multi(ctx, c):
    import rapidfuzz_amd as rfa

    qs = G.expand_list(c["queries"])
    sc = rfa.Scorer(ctx, c["metric"], ctx.strings(qs, width=1))
    got = sc.cdist(ctx.strings([G.expand(c["text"])], width=1), c["kind"], semantics=rfa.MULTI)[:, 0]
    if "expected_all" in c:
        assert np.all(np.abs(got - c["expected_all"]) <= c["tol"])
    else:
        assert list(got) == c["expected"]


@pytest.mark.parametrize("c", DATA["fixtures"], ids=[G.case_id(f) + "-" + f["file1"] for f in DATA["fixtures"]])
def test_fixture(ctx, c):
    a = G.fixture(c["file1"], c.get("sub1"))
    b = G.fixture(c["file2"], c.get("sub2"))
    assert _score(ctx, c["metric"], c["kind"], a, b, cutoff=c.get("cutoff")) == c["expected"]
