"""The C++ layer's Opcodes (rapidfuzz_amd.hpp, after details/types.hpp:357-565) on the
CPU: a small program built from the header alone converts editops to opcodes and back;
the opcodes must equal the reference's own conversion of the same editops, operation
for operation, and the round trip must give the editops back."""
import os
import subprocess

import numpy as np
import pytest

import _gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rapidfuzz-cpp_amd")

PROGRAM = r'''
#include <cstdio>
#include "rapidfuzz_amd/rapidfuzz_amd.hpp"
namespace rfa = rapidfuzz_amd;
int main()
{
    size_t n, src_len, dest_len;
    while (std::scanf("%zu %zu %zu", &n, &src_len, &dest_len) == 3) {
        rfa::Editops ops;
        ops.set_src_len(src_len);
        ops.set_dest_len(dest_len);
        for (size_t i = 0; i < n; ++i) {
            int t;
            size_t s, d;
            if (std::scanf("%d %zu %zu", &t, &s, &d) != 3) return 2;
            ops.push_back(rfa::EditOp(static_cast<rfa::EditType>(t), s, d));
        }
        rfa::Opcodes oc(ops);
        std::printf("%zu\n", oc.size());
        for (size_t i = 0; i < oc.size(); ++i)
            std::printf("%d %zu %zu %zu %zu\n", static_cast<int>(oc[i].type), oc[i].src_begin, oc[i].src_end,
                        oc[i].dest_begin, oc[i].dest_end);
        std::printf("%d\n", oc.as_editops() == ops ? 1 : 0);
    }
    return 0;
}
'''


@pytest.fixture(scope="module")
def converter(tmp_path_factory):
    d = tmp_path_factory.mktemp("opcodes")
    src = d / "conv.cpp"
    src.write_text(PROGRAM)
    exe = d / "conv"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(PKG, "include"),
                        str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(exe)


@pytest.mark.parametrize("metric", [0, 2])
def test_opcodes_equal_reference(reference, converter, metric):
    rng = np.random.default_rng(40 + metric)
    pairs = [(np.frombuffer(b"Lorem ipsum.", np.uint8).astype(np.uint32),
              np.frombuffer(b"XYZLorem ABC iPsum", np.uint8).astype(np.uint32)),
             (np.zeros(0, np.uint32), np.zeros(0, np.uint32))]
    for _ in range(200):
        s1 = rng.integers(97, 100, size=int(rng.integers(0, 40))).astype(np.uint32)
        pairs.append((s1, _gen.mutate(rng, s1, int(rng.integers(0, 10)), alpha_n=3)))
        pairs.append((s1, rng.integers(97, 100, size=int(rng.integers(0, 40))).astype(np.uint32)))
    lines, expected = [], []
    for s1, s2 in pairs:
        t, sp, dp = reference.editops(metric, s1, s2)
        lines.append(f"{len(t)} {len(s1)} {len(s2)}")
        lines += [f"{int(t[k])} {int(sp[k])} {int(dp[k])}" for k in range(len(t))]
        expected.append(reference.opcodes(metric, s1, s2))
    r = subprocess.run([converter], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = r.stdout.split("\n")
    pos = 0
    for i, exp in enumerate(expected):
        k = int(out[pos])
        got = np.array([[int(x) for x in out[pos + 1 + j].split()] for j in range(k)], dtype=np.uint64).reshape(k, 5)
        assert np.array_equal(got, exp), f"pair {i}: {got} vs {exp}"
        assert out[pos + 1 + k] == "1", f"pair {i}: round trip"
        pos += k + 2
