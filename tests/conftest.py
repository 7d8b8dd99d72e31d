import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP path")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _hip_device_count() -> int:
    try:
        import ctypes
        lib = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        if lib.hipGetDeviceCount(ctypes.byref(n)) != 0:
            return 0
        return n.value
    except OSError:
        return 0


@pytest.fixture(scope="session")
def gpu_available():
    if _hip_device_count() == 0:
        pytest.skip("no HIP device")
    return True
