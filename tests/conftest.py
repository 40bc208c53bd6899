import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


@pytest.fixture(scope="session")
def af():
    """The product module with its library loaded (GPU tests only)."""
    import arrayfire_amd

    arrayfire_amd.lib()
    n = __import__("ctypes").c_int()
    arrayfire_amd.check(arrayfire_amd.lib().af_get_device_count(__import__("ctypes").byref(n)))
    assert n.value >= 1, "no HIP device visible"
    return arrayfire_amd


@pytest.fixture
def rng():
    import numpy as np

    return np.random.default_rng(20261202)
