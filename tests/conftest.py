import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the product kernels")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    if not os.path.exists(pyoracle.LIB_PATH):
        pyoracle.build()
    return pyoracle.load()


@pytest.fixture(scope="session")
def gpu_available():
    from mcell_amd import _lib
    lib = _lib.load()  # must exist: the product path has no fallback
    return lib
