import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: full-size workload (seconds to minutes)")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The native library on a machine with a GPU; fails loudly if it is missing."""
    from pyvr_amd.hip_renderer import _lib

    L = _lib.lib()
    import ctypes

    n = ctypes.c_int()
    _lib.check(L.hvr_device_count(ctypes.byref(n)), "hvr_device_count")
    assert n.value >= 1, "no HIP device visible"
    return L
