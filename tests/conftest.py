"""Test configuration: the `gpu` marker and import paths.

`-m "not gpu"` runs the oracle pins, host logic and ABI checks on the CPU; `-m gpu` runs the parity
tests of liboal against the CPU oracle on an MI355X.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs liboal kernels)")


def gpu_available() -> bool:
    try:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.load()
    return O


@pytest.fixture(scope="session")
def gpu():
    """liboal loaded on a machine with a GPU; fails (not skips) when the library is missing."""
    if not gpu_available():
        pytest.fail("gpu test run without a visible GPU")
    from openaerolab_alpha_01_amd import _lib
    return _lib.load()
