"""The reference's xarray-output test (pyqg/tests/test_xarray_output.py:1-172)
repeated on pyqg_amd's models.  xarray is not installed here, so the test reads
the dataset's contents from `dataset_parts` (the tuples `to_dataset` hands to
xarray.Dataset) and checks that `to_dataset` raises ImportError."""
import importlib.util

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

YEAR = 24 * 60 * 60 * 360.

EXPECTED_VARS = ['q', 'u', 'v', 'ufull', 'vfull', 'p', 'dqdt', 'qh', 'uh', 'vh', 'ph', 'Ubg',
                 'Qy', 'dqhdt']
EXPECTED_DIAGS = ['EKE', 'entspec', 'APEflux', 'APEflux_div', 'KEflux', 'KEflux_div', 'APEgen',
                  'APEgenspec', 'Ensspec', 'KEspec', 'EKEdiss', 'KEspec_modal', 'PEspec_modal',
                  'APEspec', 'ENSflux', 'ENSgenspec', 'paramspec', 'paramspec_APEflux',
                  'paramspec_KEflux', 'Dissspec', 'KEfrictionspec', 'ENSDissspec',
                  'ENSfrictionspec', 'ENSparamspec']
EXPECTED_ATTRS = ['L', 'W', 'dt', 'filterfac', 'nk', 'nl', 'ntd', 'nx', 'ny', 'nz', 'rek',
                  'taveint', 'tavestart', 'tc', 'tmax', 'twrite']
EXPECTED_COORDS = ['time', 'lev', 'x', 'y', 'l', 'k']


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


def qg(pq):
    return pq.QGModel(tmax=YEAR / 2, twrite=1000, tavestart=YEAR / 3, log_level=0)


def layered(pq):
    return pq.LayeredModel(nx=64, nz=3, U=[0.05, 0.025, 0.0], V=[0., 0., 0.], L=1000.e3,
                           f=0.0001236812857687059, beta=1.2130692965249345e-11,
                           H=[500., 1750., 1750.], rho=[1025., 1025.275, 1025.640], rek=1.e-7,
                           tmax=YEAR / 2, twrite=1000, tavestart=YEAR / 3, log_level=0)


def bt(pq):
    return pq.BTModel(L=2. * np.pi, nx=256, beta=0., H=1., rek=0., rd=None, tmax=YEAR / 2,
                      twrite=1000, tavestart=YEAR / 3, dt=YEAR / 8, taveint=1, ntd=4,
                      log_level=0)


def sqg(pq):
    return pq.SQGModel(L=2. * np.pi, nx=512, tmax=YEAR / 2, beta=0., Nb=1., H=1., f_0=1.,
                       dt=YEAR / 8, taveint=1, twrite=1000, ntd=4, tavestart=YEAR / 3,
                       log_level=0)


@pytest.mark.parametrize("make", [qg, layered, bt, sqg], ids=["qg", "layered", "bt", "sqg"])
def test_dataset_contents(pq, make):
    """test_xarray_output.py:150-172: a dataset per snapshot, with the state
    variables, global attributes and coordinates; once diagnostics are filled
    every variable is a state variable or a known diagnostic, and every
    diagnostic carries the time axis."""
    from pyqg_amd.dataset import dataset_parts
    m = make(pq)
    print(f"dataset {type(m).__name__} nx={m.nx}", flush=True)
    tsnapint = YEAR / 4
    times, snaps = [], []
    for snapshot in m.run_with_snapshots(tsnapstart=m.t, tsnapint=tsnapint):
        var, coords, attrs = dataset_parts(m)
        times.append(m.t)
        snaps.append(snapshot)
        for a in EXPECTED_ATTRS:
            assert f"pyqg:{a}" in attrs, a
        for c in EXPECTED_COORDS:
            assert c in coords, c
        np.testing.assert_array_equal(coords["time"][1], [m.t])
        if snapshot < tsnapint:
            for v in EXPECTED_VARS:
                assert v in var, v
        if snapshot > tsnapint:
            for v in var:
                assert v in EXPECTED_VARS + EXPECTED_DIAGS, v
            for v in EXPECTED_DIAGS:
                if v in var:
                    assert var[v][0][0] == 'time', v
                    assert np.ndim(var[v][1]) == len(var[v][0]), v
        # the values are the model's own
        np.testing.assert_array_equal(var["q"][1][0], m.q)
        np.testing.assert_array_equal(var["qh"][1][0], m.qh)
        p = np.fft.irfftn(m.ph, axes=(-2, -1))
        assert np.abs(var["p"][1][0] - p).max() <= 1e-13 * np.abs(p).max()
    assert len(snaps) >= 2 and max(snaps) > tsnapint
    if importlib.util.find_spec("xarray") is None:
        with pytest.raises(ImportError, match="xarray"):
            m.to_dataset()
    else:
        import xarray as xr
        assert isinstance(m.to_dataset(), xr.Dataset)
