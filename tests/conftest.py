import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: takes more than ~30 s")


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


@pytest.fixture(scope="session")
def gpu_lib():
    """The HIP library, loaded; skips nothing: a GPU test without the library must fail."""
    from mcsart_amd import _lib
    lib = _lib.load()
    n = __import__("ctypes").c_int(0)
    rc = lib.mcsart_device_count(__import__("ctypes").byref(n))
    assert rc == 0 and n.value >= 1, "no HIP device visible"
    return lib
