"""Grids of 3 * 2^a points (24 ... 768): the radix-3 Stockham passes.

The transforms are checked against numpy.fft (the reference's own no-pyfftw
path), the model against the oracle and against the reference's nx=96 fixture
run (tests/golden/layered_nz3_nx96.npz), and the reference's own nx=96
diagnostic tests (pyqg/tests/test_diagnostic_tools.py:10-81,
pyqg/tests/test_diagnostics.py:155-190) are repeated on the GPU models."""
import numpy as np
import pytest

from conftest import QG_PHYS, golden, rel_l2
from oracle.qg_oracle import OracleQG

pytestmark = pytest.mark.gpu

LAYERED3 = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
                rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
                beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


@pytest.mark.parametrize("n", [24, 48, 96, 192, 384, 768])
def test_transforms_match_numpy(pq, n):
    print(f"mixed-radix transforms n={n}", flush=True)
    m = pq.QGModel(nx=n, log_level=0)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n, n))
    X = m.fft(x)
    assert rel_l2(X, np.fft.rfftn(x, axes=(-2, -1))) < 1e-14
    assert rel_l2(m.ifft(X), x) < 1e-14
    # an arbitrary half spectrum: numpy's c2r drops Im of the bins kx = 0, nx/2
    Y = rng.standard_normal(X.shape) + 1j * rng.standard_normal(X.shape)
    assert rel_l2(m.ifft(Y), np.fft.irfftn(Y, s=(n, n), axes=(-2, -1))) < 1e-14


@pytest.mark.parametrize("n", [48, 96, 192])
def test_qg_matches_oracle(pq, n):
    """100 steps through run() (fused stretches) against the oracle: the
    parity tolerance 1e-10 on q after 100 steps."""
    print(f"mixed-radix QG n={n} 100 steps", flush=True)
    q0 = 1e-7 * np.random.default_rng(7).standard_normal((2, n, n))
    m = pq.QGModel(nx=n, dt=3600.0, tmax=100 * 3600.0, log_level=0, **QG_PHYS)
    o = OracleQG(nx=n, dt=3600.0, **QG_PHYS)
    m.q = q0
    o.q = q0
    m.run()
    for _ in range(100):
        o._step_forward()
    assert m.tc == 100
    assert rel_l2(m.q, o.q) < 1e-10


def test_layered_nx96_fixture_matches_reference(pq):
    """The reference's LayeredModel at nx=96 from its fixture state: step 1
    phase by phase, then 19 more fused, against the reference's own run."""
    print("layered nx=96 fixture", flush=True)
    g = golden("layered_nz3_nx96.npz")
    m = pq.LayeredModel(nx=96, log_level=0, **LAYERED3)
    m.q = g["q0"]
    m._step_forward()
    for f in ("q", "ph", "dqhdt"):
        assert rel_l2(getattr(m, f), g["s1_" + f]) < 1e-12, f
    m.tmax = m.t + 19 * 3600.0 - 1.0
    m.run()
    assert m.tc == 20
    assert rel_l2(m.q, g["s20_q"]) < 1e-12
    assert rel_l2(m.ph, g["s20_ph"]) < 1e-12


def _fixture_models(pq):
    m1 = pq.LayeredModel(nx=96, log_level=0, **LAYERED3)
    m2 = pq.LayeredModel(nx=64, log_level=0, **LAYERED3)
    m1.q = golden("layered_nz3_nx96.npz")["q0"]
    m2.q = golden("layered_nz3_nx64.npz")["q0"]
    for m in (m1, m2):
        m._invert()
        m._calc_derived_fields()
    return m1, m2


def test_reference_diagnostic_differences(pq):
    """pyqg/tests/test_diagnostic_tools.py:10-33"""
    from pyqg_amd.diagnostic_tools import diagnostic_differences, diagnostic_similarities
    m1, m2 = _fixture_models(pq)
    diffs = diagnostic_differences(m1, m2, instantaneous=True)
    sims_hi = diagnostic_similarities(m1, m1, m2, instantaneous=True)
    sims_lo = diagnostic_similarities(m2, m1, m2, instantaneous=True)
    assert len(diffs) > 0
    for k in diffs.keys():
        assert sims_hi[k] == 1.0
        assert sims_lo[k] == 0.0


def test_reference_calc_ispec_units(pq):
    """pyqg/tests/test_diagnostic_tools.py:53-81"""
    from pyqg_amd.diagnostic_tools import calc_ispec
    for m in _fixture_models(pq):
        for a in ("q", "p"):
            for z in (0, 1):
                signal2d = getattr(m, a)[z]
                spectral = getattr(m, a + "h")[z]
                power = np.abs(spectral) ** 2 / m.M ** 2
                k, ispec = calc_ispec(m, power, averaging=False)
                dk = k[1] - k[0]
                np.testing.assert_allclose(signal2d.var(), ispec.sum() * dk, 1e-5,
                                           err_msg=f"ispec integral of {a}{z + 1}")


def test_reference_diagnostic_magnitude(pq):
    """pyqg/tests/test_diagnostics.py:155-190"""
    m1, m2 = _fixture_models(pq)
    for name in m1.diagnostics.keys():
        if name == "Dissspec":
            continue
        hi = np.abs(m1.diagnostics[name]["function"](m1)).max()
        lo = np.abs(m2.diagnostics[name]["function"](m2)).max()
        if lo == 0:
            assert hi == 0
        else:
            assert 0.33 < hi / lo < 3, f"{name} should be normalized"


def test_ensemble_with_mixed_grid_sizes(pq):
    """An ensemble holding nx=96 and nx=64 members (one bucket each) steps every
    member as its own QGModel does."""
    nxs = [96, 64, 96]
    ens = pq.Ensemble.qg(nxs, dt=3600.0, **QG_PHYS)
    q0 = [1e-7 * np.random.default_rng(30 + i).standard_normal((2, n, n))
          for i, n in enumerate(nxs)]
    ens.set_q(q0)
    ens.step(12)
    for i, n in enumerate(nxs):
        m = pq.QGModel(nx=n, dt=3600.0, tmax=12 * 3600.0, log_level=0, **QG_PHYS)
        m.q = q0[i]
        m.run()
        assert rel_l2(ens.get("q", i), m.q) < 1e-13, i
