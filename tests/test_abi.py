"""CPU checks of the boundary: the library loads and exports every declared symbol."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "rapidfuzz_amd.h")).read()
    return sorted(set(re.findall(r"\b(rfa_\w+)\s*\(", text)))


def test_header_lists_match_bindings():
    import rapidfuzz_amd as rfa

    assert sorted(rfa.EXPORTED) == declared_symbols()


def test_library_exports_all_symbols():
    import rapidfuzz_amd as rfa

    if not os.path.exists(rfa.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(rfa.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_device_is_reported_cleanly():
    """Without a GPU the context call must fail with a status, not crash."""
    import rapidfuzz_amd as rfa

    if not os.path.exists(rfa.LIB_PATH):
        pytest.skip("library not built")
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(rfa.RfaError) as e:
        rfa.Context(0)
    assert e.value.status in (3, 5)


def test_abi_version():
    import rapidfuzz_amd as rfa

    if not os.path.exists(rfa.LIB_PATH):
        pytest.skip("library not built")
    assert rfa.load_library().rfa_abi_version() == 1
