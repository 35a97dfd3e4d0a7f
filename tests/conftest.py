import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libcudadsmc.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def gpu_lib():
    """libcudadsmc.so with at least one device; GPU tests fail (not skip) without one."""
    from cudadsmc_amd import abi
    import ctypes as C
    lib = abi.load()
    n = C.c_int(0)
    abi.check(lib.dsmc_device_count(C.byref(n)))
    assert n.value >= 1, "GPU test run without a HIP device"
    return lib
