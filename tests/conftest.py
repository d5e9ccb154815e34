import ctypes
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


_GPU = None


def has_gpu():
    """Count HIP devices through the HIP runtime (no torch import)."""
    global _GPU
    if _GPU is None:
        try:
            hip = ctypes.CDLL("libamdhip64.so")
            n = ctypes.c_int(0)
            _GPU = hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
        except OSError:
            _GPU = False
    return _GPU


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.skip("no GPU visible")
    return True
