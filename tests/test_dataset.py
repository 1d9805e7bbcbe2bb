"""Dataset assembly (pyqg/xarray_output.py:118-174) on a host-only stand-in
model: names, dimensions, shapes, the real-space p and dqdt, the time axis
added to diagnostics, and the ImportError without xarray."""
import importlib.util

import numpy as np
import pytest

from pyqg_amd import dataset
from pyqg_amd.errors import DiagnosticNotFilledError


class StubModel:
    """The attributes a model exposes, as plain numpy arrays."""

    def __init__(self, nz=2, n=16):
        rng = np.random.default_rng(0)
        self.nz, self.ny, self.nx, self.nl, self.nk = nz, n, n, n, n // 2 + 1
        self.L = self.W = 1e6
        self.t, self.tc, self.dt, self.rek, self.M = 7200.0, 2, 3600.0, 1e-7, n * n
        self.x, self.y = np.meshgrid(np.arange(n) + 0.5, np.arange(n) + 10.5)
        self.kk = np.arange(self.nk) * 1.0
        self.ll = np.fft.fftfreq(n, 1.0 / n)
        for f in ("q", "u", "v", "ufull", "vfull"):
            setattr(self, f, rng.standard_normal((nz, n, n)))
        for f in ("qh", "uh", "vh", "ph", "dqhdt"):
            setattr(self, f, np.fft.rfftn(rng.standard_normal((nz, n, n)), axes=(-2, -1)))
        self.Ubg = np.array([0.025, 0.0][:nz])
        self.Qy = np.array([1e-11, 2e-11][:nz])
        self.diagnostics = {
            "EKE": {"dims": ("lev",), "description": "eke", "units": "m^2 s^-2"},
            "KEspec": {"dims": ("lev", "l", "k"), "description": "ke spectrum",
                       "units": "m^2 s^-2"},
            "EKEdiss": {"dims": ("time",), "description": "eke dissipation",
                        "units": "m^2 s^-3"},
            "Ensspec": {"dims": ("lev", "l", "k"), "description": "ens", "units": "s^-2"},
        }
        self._filled = {"EKE": np.array([1.0, 2.0]), "KEspec": np.ones((nz, n, self.nk)),
                        "EKEdiss": np.float64(3.5)}

    def ifft(self, v):
        return np.fft.irfftn(v, axes=(-2, -1))

    def get_diagnostic(self, name):
        if name not in self._filled:
            raise DiagnosticNotFilledError(name)
        return self._filled[name]


def test_variables_coordinates_attrs():
    m = StubModel()
    var, coords, attrs = dataset.dataset_parts(m)
    for name in ("q", "u", "v", "ufull", "vfull"):
        dims, data, a = var[name]
        assert dims == ("time", "lev", "y", "x") and data.shape == (1, 2, 16, 16)
        np.testing.assert_array_equal(data[0], getattr(m, name))
        assert set(a) == {"units", "long_name"}
    for name in ("qh", "uh", "vh", "ph", "dqhdt"):
        dims, data, _ = var[name]
        assert dims == ("time", "lev", "l", "k") and data.shape == (1, 2, 16, 9)
    assert var["Ubg"][0] == "lev" and var["Ubg"][1].shape == (2,)
    # p and dqdt are the inverse transforms of ph and dqhdt
    np.testing.assert_allclose(var["p"][1][0], np.fft.irfftn(m.ph, axes=(-2, -1)))
    np.testing.assert_allclose(var["dqdt"][1][0], np.fft.irfftn(m.dqhdt, axes=(-2, -1)))
    assert var["p"][2]["units"] == "m^2 s^-1"
    assert list(coords) == ["time", "lev", "lev_mid", "x", "y", "l", "k"]
    np.testing.assert_array_equal(coords["time"][1], [7200.0])
    np.testing.assert_array_equal(coords["lev"][1], [1, 2])
    np.testing.assert_array_equal(coords["lev_mid"][1], [1.5])
    np.testing.assert_array_equal(coords["x"][1], m.x[0])
    np.testing.assert_array_equal(coords["y"][1], m.y[:, 0])
    assert attrs["pyqg:nx"] == 16 and attrs["pyqg:tc"] == 2 and "pyqg:beta" not in attrs


def test_diagnostics_get_a_time_axis():
    m = StubModel()
    var, _, _ = dataset.dataset_parts(m)
    assert var["EKE"][0] == ("time", "lev") and var["EKE"][1].shape == (1, 2)
    assert var["KEspec"][0] == ("time", "lev", "l", "k")
    # a scalar diagnostic (dims ('time',) in the registry) becomes a length-one array
    assert var["EKEdiss"][0] == ("time",) and var["EKEdiss"][1].shape == (1,)
    assert var["EKEdiss"][1][0] == 3.5
    assert var["KEspec"][2] == {"long_name": "ke spectrum", "units": "m^2 s^-2"}
    assert "Ensspec" not in var          # never filled


@pytest.mark.skipif(importlib.util.find_spec("xarray") is not None, reason="xarray installed")
def test_to_dataset_needs_xarray():
    with pytest.raises(ImportError, match="xarray"):
        dataset.model_to_dataset(StubModel())
