"""The allocation-free health check against the allocating one, through the
C-ABI: two calls in a row on buckets of different sizes (the second smaller),
and the member-count limit."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_member_speed_noalloc_matches(gpu):
    import torch
    import pyqg_amd
    from pyqg_amd import _lib
    L = _lib.lib
    fn = L.qgk_member_speed_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    outs = []
    for M, n, seed in ((40, 128, 1), (7, 64, 2)):
        print(f"noalloc health check: {M} members, nx={n}", flush=True)
        ens = pyqg_amd.Ensemble.qg(np.full(M, n), dt=3600.)
        rng = np.random.default_rng(seed)
        q0 = [1e-7 * rng.standard_normal((2, n, n)) * (1 + 10 * (i % 3)) for i in range(M)]
        ens.set_q(q0)
        ens.step(5)
        b = ens.buckets[0]
        ref = b.member_speed()
        dev = torch.full((M,), -1.0, dtype=torch.float64, device="cuda")
        _lib.check(fn(b.h, ctypes.c_void_p(dev.data_ptr())))
        got = dev.cpu().numpy()
        assert np.array_equal(got, ref), (M, got, ref)
        assert np.all(ref > 0) and np.isfinite(ref).all()
        outs.append(got)
        # the scratch keeps the last call's contents: a second call on the
        # same bucket after more steps is cleared first
        ens.step(3)
        ref2 = b.member_speed()
        _lib.check(fn(b.h, ctypes.c_void_p(dev.data_ptr())))
        assert np.array_equal(dev.cpu().numpy(), ref2)
    assert outs[0].shape == (40,) and outs[1].shape == (7,)


def test_member_speed_noalloc_limit(gpu):
    import torch
    import pyqg_amd  # noqa: F401
    from pyqg_amd import _lib
    fn = _lib.lib.qgk_member_speed_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    b = _lib.Bucket(4097, 1, 8, 8, 0)
    dev = torch.zeros(4097, dtype=torch.float64, device="cuda")
    rc = fn(b.h, ctypes.c_void_p(dev.data_ptr()))   # rejected before any launch
    assert rc == _lib.E_ARG
    assert "4096" in _lib.lib.qgk_last_error().decode()
    b.close()
