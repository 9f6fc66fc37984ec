"""The oracle against the known answers of the reference's own unit tests (CPU)."""
import numpy as np
import pytest

import golden_cases as G

DATA = G.load()


def _cutoffs(c):
    kw = {}
    if "cutoff" in c:
        if isinstance(c["expected"], float) or c["kind"] >= 2 or c["metric"] in (4, 5, 6):
            kw["cutoff_d"] = c["cutoff"]
        else:
            kw["cutoff_u"] = c["cutoff"]
    return kw


@pytest.mark.parametrize("c", DATA["cases"], ids=[G.case_id(c) for c in DATA["cases"]])
def test_case(oracle, c):
    got = oracle.score(c["metric"], c["kind"], G.expand(c["s1"]), G.expand(c["s2"]),
                       weights=tuple(c.get("weights", (1, 1, 1))), **_cutoffs(c))
    if "tol" in c:
        assert abs(float(got) - c["expected"]) <= c["tol"]
    else:
        assert got == c["expected"]


@pytest.mark.parametrize("c", DATA["multi"], ids=[G.case_id(c) for c in DATA["multi"]])
def test_multi(oracle, c):
    qs = G.expand_list(c["queries"])
    got = oracle.cdist(c["metric"], c["kind"], qs, [G.expand(c["text"])], multi=True)[:, 0]
    if "expected_all" in c:
        assert np.all(np.abs(got - c["expected_all"]) <= c["tol"])
    else:
        assert list(got) == c["expected"]


@pytest.mark.parametrize("c", [f for f in DATA["fixtures"] if "issue9" in f["file1"]],
                         ids=lambda f: G.case_id(f))
def test_fixture_issue9(oracle, c):
    a = G.fixture(c["file1"], c.get("sub1"))
    b = G.fixture(c["file2"], c.get("sub2"))
    kw = {"cutoff_u": c["cutoff"]} if "cutoff" in c else {}
    assert oracle.score(c["metric"], c["kind"], a, b, **kw) == c["expected"]
