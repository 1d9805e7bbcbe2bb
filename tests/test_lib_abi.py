"""The C-ABI library loads and exports every entry point include/pyqg_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "pyqg_amd.h")
LIB = os.path.join(ROOT, "pyqg_amd", "_qgk.so")
CNN_HEADER = os.path.join(ROOT, "include", "pyqg_amd_cnn.h")
CNN_LIB = os.path.join(ROOT, "pyqg_amd", "_qgcnn.so")


def declared(header=HEADER, prefix="qgk_"):
    text = open(header).read()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*|void\*)\s+(" + prefix +
                                 r"\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "qgk_create" in names and "qgk_step" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libqgk is not built: run python -m pyqg_amd.build")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from pyqg_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()


def test_create_fails_loudly_without_device():
    """With no usable HIP device the product raises instead of falling back."""
    from pyqg_amd import _lib
    n = ctypes.c_int(0)
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0:
            pytest.skip("a GPU is present")
    except OSError:
        pass
    with pytest.raises(_lib.QGKError):
        _lib.Bucket(1, 2, 32, 32, _lib.FULL_STATE)


def test_bad_sizes_rejected_before_device_use():
    from pyqg_amd import _lib
    # sizes are 2^a in [8, 1024] or 3 * 2^a in [24, 768]
    for n in (40, 12, 1536, 2048, 100):
        with pytest.raises(ValueError):
            _lib.Bucket(1, 2, n, n, 0)
    with pytest.raises(ValueError):
        _lib.Bucket(0, 2, 32, 32, 0)


def test_cnn_library_exports_every_declared_symbol():
    if not os.path.exists(CNN_LIB):
        pytest.fail("libqgcnn is not built: run python pyqg_amd/build.py")
    lib = ctypes.CDLL(CNN_LIB)
    names = declared(CNN_HEADER, "qgcnn_")
    assert len(names) >= 9
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from pyqg_amd.closure import cnn
    assert sorted(cnn.EXPORTS) == names
