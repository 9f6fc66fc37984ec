import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "rapidfuzz-cpp_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle_py import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from oracle_py import Reference, reference_available

    if not reference_available():
        pytest.skip("oracle/_ref/librfref.so not built (needs the reference sources)")
    return Reference()


@pytest.fixture(scope="session")
def ctx():
    import rapidfuzz_amd as rfa

    c = rfa.Context(0)
    yield c
    c.close()
