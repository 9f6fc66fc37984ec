"""The C++ layer's Editops / Opcodes algebra (slice, remove_slice, reverse, inverse,
remove_subsequence; Opcodes slice / inverse / reverse and the conversions; after
details/types.hpp:108-565) on the CPU: tests/cpp/test_editops_algebra.cpp, built from the
header alone, runs each case; the reference library runs the same case; results and the
cases that throw must agree.  Inputs are the reference's own editops of random pairs."""
import os
import subprocess

import numpy as np
import pytest

import _gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rapidfuzz-cpp_amd")


@pytest.fixture(scope="module")
def algebra(tmp_path_factory):
    d = tmp_path_factory.mktemp("algebra")
    exe = d / "algebra"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(PKG, "include"),
                        os.path.join(ROOT, "tests", "cpp", "test_editops_algebra.cpp"), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


def _cases(reference, rng):
    """(op, a, b, c, src_len, dest_len, ops, sub) tuples over editops of random pairs."""
    cases = []
    for it in range(160):
        s1 = rng.integers(97, 101, size=int(rng.integers(0, 30))).astype(np.uint32)
        s2 = _gen.mutate(rng, s1, int(rng.integers(0, 8)), alpha_n=4) if it % 3 else \
            rng.integers(97, 101, size=int(rng.integers(0, 30))).astype(np.uint32)
        t, sp, dp = reference.editops(0 if it % 2 else 2, s1, s2)
        ops = np.stack([t.astype(np.int64), sp.astype(np.int64), dp.astype(np.int64)], axis=1) if len(t) else \
            np.zeros((0, 3), np.int64)
        n = len(ops)
        for op in range(9):
            if op in (0, 1, 5, 8):
                for _ in range(4):
                    a = int(rng.integers(-n - 3, n + 4))
                    b = int(rng.integers(-n - 3, n + 4))
                    c = int(rng.choice([1, 1, 2, 3, 0, -1]))
                    cases.append((op, a, b, c, len(s1), len(s2), ops, None))
            elif op == 4:
                # a true subsequence, and one longer than the list (the reference throws before
                # writing; for an element that is not in the list it writes past its result)
                keep = rng.random(n) < 0.5
                cases.append((op, 0, 0, 1, len(s1), len(s2), ops, ops[keep]))
                cases.append((op, 0, 0, 1, len(s1), len(s2), ops, np.concatenate([ops, ops[:1]])))
            else:
                cases.append((op, 0, 0, 1, len(s1), len(s2), ops, None))
    return cases


def _line(case):
    op, a, b, c, sl, dl, ops, sub = case
    sub = np.zeros((0, 3), np.int64) if sub is None else sub
    parts = [op, a, b, c, sl, dl, len(ops)] + [int(x) for x in ops.ravel()] + [len(sub)] + [int(x) for x in sub.ravel()]
    return " ".join(str(x) for x in parts)


def test_algebra_equals_reference(reference, algebra):
    rng = np.random.default_rng(77)
    cases = _cases(reference, rng)
    r = subprocess.run([algebra], input="\n".join(_line(c) for c in cases) + "\n", capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout.strip("\n").split("\n")
    assert len(out) == len(cases)
    errs = 0
    for case, got in zip(cases, out):
        op, a, b, c, sl, dl, ops, sub = case
        exp = reference.editops_algebra(op, ops, sl, dl, a, b, c, sub)
        if exp is None:
            assert got == "err", case
            errs += 1
            continue
        vals = [int(x) for x in got.split()]
        esl, edl, rows = exp
        assert vals[:3] == [esl, edl, len(rows)], (case, got)
        assert vals[3:] == [int(x) for x in rows.ravel()], (case, got)
    assert 0 < errs < len(cases)


def test_inverse_and_subsequence_properties(reference, algebra):
    """inverse twice is the identity; removing the whole list leaves nothing; an element that
    is not in the list is rejected."""
    rng = np.random.default_rng(78)
    cases = []
    for _ in range(100):
        s1 = rng.integers(97, 100, size=int(rng.integers(0, 25))).astype(np.uint32)
        s2 = _gen.mutate(rng, s1, int(rng.integers(0, 6)), alpha_n=3)
        t, sp, dp = reference.editops(0, s1, s2)
        ops = np.stack([t.astype(np.int64), sp.astype(np.int64), dp.astype(np.int64)], axis=1) if len(t) else \
            np.zeros((0, 3), np.int64)
        inv = reference.editops_algebra(3, ops, len(s1), len(s2))
        cases.append((3, 0, 0, 1, len(s2), len(s1), inv[2], None))
        cases.append((4, 0, 0, 1, len(s1), len(s2), ops, ops))
    bad = np.array([[1, 5, 5], [2, 0, 0]], np.int64)
    cases.append((4, 0, 0, 1, 10, 10, np.array([[1, 5, 5], [3, 6, 5]], np.int64), bad))
    r = subprocess.run([algebra], input="\n".join(_line(c) for c in cases) + "\n", capture_output=True, text=True,
                       timeout=60)
    out = r.stdout.strip("\n").split("\n")
    assert out[-1] == "err"
    for i in range(0, len(cases) - 1, 2):
        ops = cases[i + 1][6]
        back = [int(x) for x in out[i].split()]
        assert back[3:] == [int(x) for x in ops.ravel()]
        empty = [int(x) for x in out[i + 1].split()]
        assert empty[2] == 0
