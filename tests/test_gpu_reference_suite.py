"""The reference's own model-construction tests, repeated on pyqg_amd's models:
the layered stretching matrix and background gradients
(pyqg/tests/test_layered.py:9-58), vertical modes and the modal projection
(pyqg/tests/test_vertical_modes.py:5-65) and describe_diagnostics
(pyqg/tests/test_diagnostics.py:10-14)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


@pytest.fixture(scope="module")
def layered(pq):
    m = pq.LayeredModel(nz=3, U=[.1, .05, .0], V=[.1, .05, .0], rho=[.1, .3, .5],
                        H=[.1, .1, .3], f=1., beta=0., log_level=0)
    F11 = m.f2 / m.gpi[0] / m.Hi[0]
    F12 = m.f2 / m.gpi[0] / m.Hi[1]
    F22 = m.f2 / m.gpi[1] / m.Hi[1]
    F23 = m.f2 / m.gpi[1] / m.Hi[2]
    S = np.zeros((3, 3))
    S[0, 0], S[0, 1] = -F11, F11
    S[1, 0], S[1, 1], S[1, 2] = F12, -(F12 + F22), F22
    S[2, 1], S[2, 2] = F23, -F23
    return m, S


def test_stretching(layered):
    """pyqg/tests/test_layered.py:32-46"""
    m, S = layered
    assert np.all(m.S.sum(axis=1) == 0.), "zero is not an eigenvalue of S"
    HS = np.dot(np.diag(m.Hi), m.S)
    np.testing.assert_allclose(HS, HS.T, atol=1e-16, err_msg="Hi*S is not symmetric")
    np.testing.assert_allclose(m.S, S, atol=1e-16, err_msg="unmatched stretching matrix")


def test_init_background(layered):
    """pyqg/tests/test_layered.py:48-58"""
    m, S = layered
    np.testing.assert_allclose(-np.einsum('ij,j->i', S, m.Ubg), m.Qy, atol=1e-16)
    np.testing.assert_allclose(np.einsum('ij,j->i', S, m.Vbg), m.Qx, atol=1e-16)


@pytest.fixture(scope="module")
def modes(pq):
    m = pq.LayeredModel(nz=2, U=[.1, .05], V=[.1, .0], H=[1., 1.], delta=1., rd=5., f=1.,
                        beta=0., log_level=0)
    m.set_q(np.random.default_rng(0).standard_normal((m.nz, m.ny, m.nx)))
    m.vertical_modes()
    return m


def test_vertical_mode_radii(modes):
    """pyqg/tests/test_vertical_modes.py:25-32"""
    m = modes
    radii = np.array([np.sqrt(m.g * m.H) / np.abs(m.f), m.rd])
    np.testing.assert_allclose(m.radii, radii, atol=1e-16, err_msg="wrong deformation radii")


def test_vertical_modes_normalized(modes):
    """pyqg/tests/test_vertical_modes.py:34-53"""
    m = modes
    p = (m.Hi[:, np.newaxis] * m.pmodes) / m.H
    np.testing.assert_allclose(p[0, 0], p[1, 0], atol=1e-16)      # barotropic: depth invariant
    np.testing.assert_allclose(p[0, 1] + p[1, 1], 0., atol=1e-16)  # baroclinic: zero integral
    np.testing.assert_allclose(p[0, 0] + p[1, 0], 1., atol=1e-16)  # normalization


def test_modal_projection_round_trip(modes):
    """pyqg/tests/test_vertical_modes.py:56-62"""
    m = modes
    qn = m.modal_projection(m.qh, forward=True)
    qh = m.modal_projection(qn, forward=False)
    np.testing.assert_allclose(m.qh, qh, atol=1e-16)


def test_describe_diagnostics(pq, capsys):
    """pyqg/tests/test_diagnostics.py:10-14: QGModel(1) (beta = 1, positional)."""
    m = pq.QGModel(1, log_level=0)
    assert m.beta == 1
    m.describe_diagnostics()
    assert "KEspec" in capsys.readouterr().out
