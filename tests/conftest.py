import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


def golden(name):
    """Load one committed golden-vector file (plain arrays, no pickles)."""
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


QG_PHYS = dict(L=1e6, H1=500.0, delta=0.25, U1=0.025, U2=0.0, beta=1.5e-11, rd=15000.0,
               rek=5.787e-7, filterfac=23.6)


def rel_l2(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


@pytest.fixture(scope="session")
def gpu():
    """Skip unless a HIP device is usable; GPU tests are marked `gpu` anyway."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        if hip.hipGetDeviceCount(ctypes.byref(n)) != 0 or n.value == 0:
            pytest.skip("no HIP device")
    except OSError:
        pytest.skip("no HIP runtime")
    return True
