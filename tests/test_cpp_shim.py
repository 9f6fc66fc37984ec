"""The C++ drop-in classes (rapidfuzz-cpp_amd/include/rapidfuzz_amd/rapidfuzz_amd.hpp)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rapidfuzz-cpp_amd")
BIN = os.path.join(PKG, "lib", "test_shim")


def test_header_compiles_standalone(tmp_path):
    src = tmp_path / "inc.cpp"
    src.write_text('#include "rapidfuzz_amd/rapidfuzz_amd.hpp"\n'
                   'int main() { rapidfuzz_amd::experimental::MultiIndel<32> m(3); return (int)m.result_count() - 8; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I" + os.path.join(PKG, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _run():
    if not os.path.exists(BIN):
        pytest.skip("test_shim not built (make -C rapidfuzz-cpp_amd)")
    return subprocess.run([BIN], capture_output=True, text=True, timeout=300)


def test_without_device_fails_cleanly():
    """On a machine without a GPU the classes throw Error(RFA_ERR_NO_DEVICE), nothing crashes."""
    import rapidfuzz_amd as rfa

    try:
        rfa.Context(0).close()
        pytest.skip("a GPU is present")
    except rfa.RfaError as e:
        assert e.status == 5, e  # RFA_ERR_NO_DEVICE
    r = _run()
    assert r.returncode == 77, r.stdout + r.stderr


@pytest.mark.gpu
def test_shim_on_gpu():
    r = _run()
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("OK")
