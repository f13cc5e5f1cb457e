import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libhydrofv_amd.so kernels)")
    config.addinivalue_line("markers", "slow: long-running (large grids)")


@pytest.fixture(scope="session")
def gpulib():
    """libhydrofv_amd.so with prototypes attached.  A missing library is an
    error, not a skip: GPU tests must run the HIP path."""
    import hydrofv_amd
    return hydrofv_amd.load()


@pytest.fixture(scope="session")
def reflib():
    """The reference library compiled from its sources (oracle/_ref), if built."""
    import ctypes
    import _oracle as O
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    from hydrofv_amd import abi
    return abi.bind(ctypes.CDLL(O.REF_SO))
