"""Pins the CPU oracle (oracle/qg_oracle.py) before it is trusted as the checker.

* golden vectors written from the reference build (tests/golden/make_golden.py);
* the regression values the reference's own tests hold
  (pyqg/tests/test_reference_solns.py);
* closed forms restated from pyqg/tests/test_model.py, test_advection.py,
  test_fft2.py, test_layered.py;
* a brute-force DFT for the transforms.
"""
import numpy as np
import pytest

from conftest import QG_PHYS, golden, rel_l2
from oracle.qg_oracle import (OracleBT, OracleLayered, OracleQG, OracleSQG, ab_coefficients)

STATE = ["q", "qh", "ph", "u", "v", "uq", "vq", "uqh", "vqh", "dqhdt", "dqhdt_p", "dqhdt_pp"]


def close(a, b, rtol=1e-13):
    a, b = np.asarray(a), np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    assert np.abs(a - b).max() <= rtol * scale, (np.abs(a - b).max(), scale)


# ----------------------------------------------------------------- golden
def test_qg_startup_all_fields_match_reference():
    g = golden("qg_nx32_startup.npz")
    m = OracleQG(nx=32, dt=7200.0, **QG_PHYS)
    for f in ["kk", "ll", "filtr", "Qy", "Ubg", "_k2l2"]:
        close(getattr(m, f), g["const_" + f], 0)
    close(m.a.real, g["const_a"], 0)
    m.q = g["q0"]
    close(m.qh, g["qh0"], 0)
    for step in range(1, 4):
        m._step_forward()
        if step in (1, 3):
            for f in STATE:
                close(getattr(m, f), g[f"s{step}_{f}"], 1e-14)
            assert m.t == g[f"s{step}_t"] and m.tc == g[f"s{step}_tc"]
            assert m.ablevel == g[f"s{step}_ablevel"]


def test_qg_config1_trajectory():
    g = golden("qg_nx64_cfg1.npz")
    m = OracleQG(nx=64, dt=7200.0, **QG_PHYS)
    m.q = g["q0"]
    for step in range(1, 1001):
        m._step_forward()
        if step in (100, 1000):
            assert rel_l2(m.q, g[f"s{step}_q"]) < 1e-13
            assert m.t == g[f"s{step}_t"]


LAYERED3 = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
                rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
                beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)


def test_layered_nz3_from_spun_up_fixture():
    g = golden("layered_nz3_nx64.npz")
    m = OracleLayered(nx=64, **LAYERED3)
    close(m.S, g["const_S"], 1e-15)
    close(m.Qy, g["const_Qy"], 1e-15)
    for key in ("5_5", "3_7", "m2_1"):
        i, j = [int(x.replace("m", "-")) for x in key.split("_")]
        close(m.a[:, :, i, j].real, g["const_a_" + key], 1e-13)
    m.q = g["q0"]
    m._step_forward()
    for f in ("q", "ph", "dqhdt"):
        close(getattr(m, f), g["s1_" + f], 1e-12)
    for _ in range(19):
        m._step_forward()
    assert rel_l2(m.q, g["s20_q"]) < 1e-12
    assert rel_l2(m.ph, g["s20_ph"]) < 1e-12


def test_layered_nz3_nx96_from_spun_up_fixture():
    """The reference's nx=96 fixture state (a 3 * 2^5 grid)."""
    g = golden("layered_nz3_nx96.npz")
    m = OracleLayered(nx=96, **LAYERED3)
    m.q = g["q0"]
    m._step_forward()
    for f in ("q", "ph", "dqhdt"):
        close(getattr(m, f), g["s1_" + f], 1e-12)
    for _ in range(19):
        m._step_forward()
    assert rel_l2(m.q, g["s20_q"]) < 1e-12
    assert rel_l2(m.ph, g["s20_ph"]) < 1e-12


def test_layered_nz10():
    g = golden("layered_nz10_nx32.npz")
    m = OracleLayered(nz=10, nx=32, H=g["in_H"], rho=g["in_rho"], U=g["in_U"],
                      V=np.zeros(10), L=2e6, beta=1.5e-11, rek=5.787e-7, dt=3600.0,
                      f=0.0001236812857687059, g=9.81)
    close(m.a[:, :, 5, 5].real, g["const_a_5_5"], 1e-12)
    close(m.a[:, :, -2, 1].real, g["const_a_m2_1"], 1e-12)
    m.q = g["q0"]
    for _ in range(10):
        m._step_forward()
    assert rel_l2(m.q, g["s10_q"]) < 1e-11


def test_bt_sqg_100_steps():
    g = golden("bt_sqg_nx64.npz")
    m = OracleBT(L=2. * np.pi, nx=64, beta=0., H=1., rek=0., rd=None, dt=0.0025)
    close(m.a.real, g["bt_a"], 0)
    close(m.filtr, g["bt_filtr"], 0)
    m.q = g["bt_q0"]
    for _ in range(100):
        m._step_forward()
    assert rel_l2(m.q, g["bt_s100_q"]) < 1e-13
    assert m.t == g["bt_s100_t"]
    m = OracleSQG(L=2. * np.pi, nx=64, beta=0., H=1., dt=0.0025)
    close(m.a.real, g["sqg_a"], 0)
    m.q = g["sqg_q0"]
    for _ in range(100):
        m._step_forward()
    assert rel_l2(m.q, g["sqg_s100_q"]) < 1e-13


def q_param(m):
    return -2e-7 * m.q + 1e-13 * np.cos(2 * np.pi * m.x / m.L)[np.newaxis]


def uv_param_curl(m):
    return 1e-6 * m.v, 1e-6 * m.u + 1e-9 * np.cos(2 * np.pi * m.x / m.L)[np.newaxis]


def uv_param(m):
    return 1e-6 * m.v, -1e-6 * m.u + 1e-9 * np.sin(2 * np.pi * m.y / m.W)[np.newaxis]


def test_parameterized_steps():
    g = golden("qg_param_nx32.npz")
    m = OracleQG(nx=32, dt=7200.0, q_parameterization=q_param, uv_parameterization=uv_param,
                 **QG_PHYS)
    m.q = g["q0"]
    for _ in range(3):
        m._step_forward()
    for f in ("q", "qh", "dqhdt", "dqh", "duh", "dvh"):
        close(getattr(m, f), g["s3_" + f], 1e-13)


# ----------------------------------------------------- long regression runs
def _gaussian(m):
    return (np.exp(-(2. * (m.x - 1.75 * np.pi / 2)) ** 2. - (2. * (m.y - np.pi)) ** 2)
            + np.exp(-(2. * (m.x - 2.25 * np.pi / 2)) ** 2. - (2. * (m.y - np.pi)) ** 2))


@pytest.mark.parametrize("kind", ["bt", "sqg"])
def test_long_runs_match_reference_values(kind):
    """test_reference_solns.py:97-175: 2000 steps; values held by the reference
    (rtol 1e-5 / 1e-4) and the reference build's own values (tighter)."""
    lr = golden("long_runs.npz")
    if kind == "bt":
        m = OracleBT(L=2. * np.pi, nx=64, tmax=5, beta=0., H=1., rek=0., rd=None, dt=0.0025)
        held = dict(qnorm=89101.741238768518, pnorm=1493.217664248918, ke=0.9950360837282386)
        rtol = 1e-5
    else:
        m = OracleSQG(L=2. * np.pi, nx=64, tmax=5., beta=0., H=1., dt=0.0025)
        held = dict(qnorm=7847.5169609881032, pnorm=1346.8874163575524, ke=0.9579493831013508)
        rtol = 1e-4
    ph = m.fft(_gaussian(m)[np.newaxis])
    KEaux = m.spec_var(m.filtr * m.wv * ph) / 2.
    ph = ph / np.sqrt(KEaux)
    m.q = m.ifft(-m.wv2 * ph if kind == "bt" else m.wv * ph)
    m.run()
    assert m.t == 5.000000000000082
    got = dict(qnorm=(m.q ** 2).sum(), pnorm=(m.ifft(m.ph) ** 2).sum(), ke=m._calc_ke())
    for k, v in held.items():
        np.testing.assert_allclose(got[k], v, rtol=rtol)
        np.testing.assert_allclose(got[k], lr[f"{kind}_{k}"], rtol=1e-11)


def test_two_layer_regression():
    """test_reference_solns.py:9-51 (q1 norm after 3 model years, rtol 1e-2)."""
    lr = golden("long_runs.npz")
    year = 360 * 86400.
    m = OracleQG(nx=32, L=1e6, beta=1.5e-11, rek=5.787e-7, rd=30000.0, delta=0.25, U1=0.05,
                 U2=0.0, filterfac=18.4, dt=12800., tmax=3 * year)
    m.set_q1q2(1e-6 * np.cos(2 * 5 * np.pi * m.x / m.L) + 1e-7 * np.cos(2 * 5 * np.pi * m.y / m.W),
               np.zeros_like(m.x))
    m.run()
    assert m.t == 93312000.0
    q1norm = (m.q[0] ** 2).sum()
    np.testing.assert_allclose(q1norm, 9.561430503712755e-08, rtol=1e-2)
    np.testing.assert_allclose(q1norm, lr["two_layer_q1norm"], rtol=1e-9)


# ------------------------------------------------------------ closed forms
def _closed_form_model():
    return OracleQG(beta=0., U1=0., U2=0., filterfac=0.)


def test_fft_bin_amplitudes():
    """test_model.py:16-65"""
    m = _closed_form_model()
    amp = (m.nx / 2) ** 2 * m.ny ** 2
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            m.set_q1q2(np.cos(k * m.x), np.sin(l * m.y))
            a = np.real(m.qh * m.qh.conj())
            np.testing.assert_allclose(a[0, 0, kw], amp, 1e-15)
            np.testing.assert_allclose(a[1, lw, 0], amp, 1e-15)
            np.testing.assert_allclose(a[1, -lw, 0], amp, 1e-15)
            a[0, 0, kw] = a[1, lw, 0] = a[1, -lw, 0] = 0
            assert np.abs(a).max() <= 1e-15 * amp ** 0 + 1e-15


def test_inversion_barotropic():
    """test_model.py:67-94"""
    m = _closed_form_model()
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            q = np.cos(k * m.x) + np.sin(l * m.y)
            m.set_q1q2(q, q)
            m._invert()
            for z in range(2):
                np.testing.assert_allclose(m.u[z], l ** -1 * np.cos(l * m.y), 1e-13)
                np.testing.assert_allclose(m.v[z], k ** -1 * np.sin(k * m.x), 1e-13)


def test_advection_closed_form():
    """test_model.py:108-158: energy only at 2k and 2l."""
    m = _closed_form_model()
    amp = (m.nx / 2) ** 2 * m.ny ** 2
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            m.set_q1q2(np.cos(k * m.x), np.sin(l * m.y))
            m._u = np.stack([np.sin(k * m.x), np.zeros_like(m.x)])
            m._v = np.stack([np.zeros_like(m.y), np.cos(l * m.y)])
            m.Ubg = np.array([0., 0.])
            m._do_advection()
            t = np.real(m.dqhdt * m.dqhdt.conj())
            np.testing.assert_allclose(t[0, 0, 2 * kw], k ** 2 * amp, 1e-14)
            np.testing.assert_allclose(t[1, 2 * lw, 0], l ** 2 * amp, 1e-14)
            np.testing.assert_allclose(t[1, -2 * lw, 0], l ** 2 * amp, 1e-14)
            t[0, 0, 2 * kw] = t[1, 2 * lw, 0] = t[1, -2 * lw, 0] = 0
            assert np.abs(t).max() <= 1e-14


def test_friction_closed_form():
    """test_model.py:160-183"""
    m = _closed_form_model()
    m.set_q1q2(np.cos(2 * np.pi * 3 / m.L * m.x), np.sin(2 * np.pi * 2 / m.W * m.y))
    m._invert()
    m._dqhdt[:] = 0.
    m._do_friction()
    np.testing.assert_allclose(m.dqhdt[-1], m.rek * m.wv2 * m.ph[-1], 1e-15)
    assert np.all(m.dqhdt[:-1] == 0)


def test_timestepping_startup():
    """test_model.py:314-340: AB1, AB2, AB3 with a constant tendency."""
    m = _closed_form_model()
    m.set_q1q2(np.zeros_like(m.x), np.zeros_like(m.x))
    rng = np.random.default_rng(5)
    d = rng.random(m.dqhdt.shape) + 1j * rng.random(m.dqhdt.shape)
    m._dqhdt[:] = d
    for n in (1, 2, 3):
        m._forward_timestep()
        np.testing.assert_allclose(m.qh, n * m.dt * d)


def test_plane_wave_jacobian_vanishes():
    """test_advection.py:4-52 (plane-wave Jacobian, nx=256)."""
    m = OracleQG(L=2. * np.pi, nx=256, U1=0)
    for kx, ky in [(1., 2.), (5., 0.), (10., 10.), (0., 21.), (24., 12.), (2., 49.)]:
        m.set_q1q2(np.cos(kx * m.x + ky * m.y), np.zeros_like(m.x))
        m._invert()
        uq, vq = m.u * m.q, m.v * m.q
        jac = m.ifft(1j * m.k * m.fft(uq) + 1j * m.l * m.fft(vq))
        assert np.abs(jac).sum() * m.dx * m.dy / m.L ** 2 < 1e-13


def test_ab_coefficients():
    assert ab_coefficients(0, 2.0) == (2.0, 0.0, 0.0)
    assert ab_coefficients(1, 2.0) == (3.0, -1.0, 0.0)
    assert ab_coefficients(5, 12.0) == (23.0, -16.0, 5.0)


# ------------------------------------------------------- transforms pinned
def test_transforms_against_brute_force_dft():
    rng = np.random.default_rng(0)
    nz, n = 2, 8
    x = rng.standard_normal((nz, n, n))
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    full = np.einsum("ab,zbc,cd->zad", W, x, W)
    m = OracleQG(nx=n, **QG_PHYS)
    close(m.fft(x), full[:, :, : n // 2 + 1], 1e-14)
    # c2r discards the imaginary parts of the kx = 0 and kx = n/2 columns
    X = rng.standard_normal((nz, n, n // 2 + 1)) + 1j * rng.standard_normal((nz, n, n // 2 + 1))
    Y = np.fft.ifft(X, axis=-2)
    Y[..., 0] = Y[..., 0].real
    Y[..., -1] = Y[..., -1].real
    ext = np.concatenate([Y, np.conj(Y[..., 1:-1][..., ::-1])], axis=-1)
    brute = np.einsum("zak,kb->zab", ext, np.conj(W)).real / n
    close(m.ifft(X), brute, 1e-14)


def test_parseval():
    """test_fft2.py:8-34"""
    m = OracleQG(L=2. * np.pi, beta=0., U1=0., U2=0., filterfac=0.)
    rng = np.random.default_rng(1)
    p = rng.standard_normal((m.nz, m.ny, m.nx))
    p -= p.mean(axis=(-2, -1))[:, np.newaxis, np.newaxis]
    ph = m.fft(p)
    var_spec = 2. * np.abs(ph) ** 2 / m.M ** 2
    var_spec[..., 0] /= 2
    var_spec[..., -1] /= 2
    assert np.abs((p.var(axis=(-2, -1)) - var_spec.sum(axis=(-2, -1)))
                  / p.var(axis=(-2, -1))).sum() < 1e-14


def test_layered_inversion_is_matrix_inverse():
    """a(k,l) (S - kappa^2 I) = I away from kappa = 0; the 3x3 cofactor check of
    test_layered.py:60-84."""
    m = OracleLayered(nx=16, **{k: v for k, v in LAYERED3.items() if k != "nz"}, nz=3)
    for (j, i) in [(5, 5), (1, 0), (3, 7), (-2, 1)]:
        M = m.S - np.eye(3) * m.wv2[j, i]
        np.testing.assert_allclose(m.a[:, :, j, i].real @ M, np.eye(3), atol=1e-12)
    M = m.S - np.eye(3) * m.wv2[5, 5]
    cof = np.array([[np.linalg.det(np.delete(np.delete(M, r, 0), c, 1)) * (-1) ** (r + c)
                     for c in range(3)] for r in range(3)])
    np.testing.assert_allclose(m.a[:, :, 5, 5].real, cof.T / np.linalg.det(M), rtol=1e-12)
    assert np.all(m.a[:, :, 0, 0] == 0)


def test_planted_mistake_is_caught():
    """A wrong sign in the friction term must break the golden comparison."""
    g = golden("qg_nx32_startup.npz")
    m = OracleQG(nx=32, dt=7200.0, **QG_PHYS)
    m.q = g["q0"]
    m.rek = -m.rek
    m._step_forward()
    assert np.abs(m.dqhdt - g["s1_dqhdt"]).max() > 1e-10 * np.abs(g["s1_dqhdt"]).max()


# ------------------------------------------------------------------ diagnostics
DIAG_CASES = {
    # the configurations of tests/golden/make_golden.py DIAG_CASES
    "qg": (OracleQG, dict(nx=32, dt=7200.0, **QG_PHYS), 0.0, 7200.0, 5, True, True, 12),
    "qg_sparse": (OracleQG, dict(nx=32, dt=3600.0, **QG_PHYS), 5 * 3600.0, 3 * 3600.0, 7,
                  False, False, 20),
    "layered3": (OracleLayered, dict(nx=32, **LAYERED3), 0.0, 3600.0, 6, False, True, 12),
    "layered2": (OracleLayered, dict(nx=32, nz=2, rd=15000.0, delta=0.25, H=[500., 2000.],
                                     U=[0.025, 0.0], V=[0.0, 0.0], rho=[1025.0, 1025.3],
                                     dt=3600.0), 0.0, 3600.0, 8, True, False, 8),
    "bt": (OracleBT, dict(nx=32, beta=1.5e-11, rd=15000.0, H=1.0, U=0.01, dt=3600.0),
           0.0, 3600.0, 9, False, False, 8),
    "sqg": (OracleSQG, dict(nx=32, beta=0.0, H=1.0, U=0.0, dt=3600.0), 0.0, 3600.0, 10,
            False, False, 8),
}


def diag_case(case):
    """Oracle model of one golden diagnostics case, stepped with its accumulator."""
    from oracle.qg_diagnostics import Accumulator, step_with_diagnostics
    cls, kw, tavestart, taveint, seed, qp, uvp, steps = DIAG_CASES[case]
    extra = {}
    if qp:
        extra["q_parameterization"] = q_param
    if uvp:
        extra["uv_parameterization"] = uv_param_curl
    m = cls(**kw, **extra)
    m.q = 1e-6 * np.random.default_rng(seed).standard_normal((m.nz, m.ny, m.nx))
    acc = Accumulator(m, tavestart=tavestart, taveint=taveint)
    for _ in range(steps):
        step_with_diagnostics(m, acc)
    return m, acc


def diag_close(a, b, tol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    scale = np.abs(b).max()
    if scale == 0.0:
        return np.abs(a).max() == 0.0
    return np.abs(a - b).max() <= tol * scale


@pytest.mark.parametrize("case", list(DIAG_CASES))
def test_diagnostics_match_reference(case):
    """Every registered diagnostic's time mean and count equal the reference
    build's (tests/golden/diagnostics.npz)."""
    from oracle.qg_diagnostics import names
    g = golden("diagnostics.npz")
    m, acc = diag_case(case)
    assert rel_l2(m.q, g[f"{case}_q_end"]) < 1e-12
    expected = sorted(k[len(case) + 7:] for k in g.files if k.startswith(f"{case}_count_"))
    assert sorted(names(m)) == expected
    for n in expected:
        assert acc.count == int(g[f"{case}_count_{n}"]), n
        assert diag_close(acc.get(n), g[f"{case}_{n}"], 1e-11), n


def test_diagnostics_planted_mistake_is_caught():
    """Dissspec with the AB history terms dropped (forward-Euler weights at
    every step) breaks the golden comparison."""
    import oracle.qg_diagnostics as D
    g = golden("diagnostics.npz")
    m, acc = diag_case("qg")
    good = acc.get("Dissspec")
    assert diag_close(good, g["qg_Dissspec"], 1e-11)
    orig = D._dissipation_spectrum

    def euler_only(mm):
        return np.stack([(mm.filtr - 1.0) * (mm.qh[k] + mm.dt * mm.dqhdt[k])
                         for k in range(mm.nz)])
    try:
        D._dissipation_spectrum = euler_only
        m, acc = diag_case("qg")
        assert not diag_close(acc.get("Dissspec"), g["qg_Dissspec"], 1e-6)
    finally:
        D._dissipation_spectrum = orig


def test_dissspec_identity():
    """test_diagnostics.py:101-154: Dissspec equals the filtered minus the
    unfiltered contribution of the AB update to the energy."""
    from oracle.qg_diagnostics import evaluate, attach_constants, derived_fields
    m = OracleQG(nx=32, dt=3600.0, **QG_PHYS)
    m.q = 1e-6 * np.random.default_rng(0).standard_normal((2, 32, 32))
    for _ in range(30):
        m._step_forward()
    m._invert()
    m._do_advection()
    m._do_friction()
    attach_constants(m)
    derived_fields(m)
    dt1, dt2, dt3 = ab_coefficients(m.ablevel, m.dt)
    rhs = m.qh + dt1 * m.dqhdt + dt2 * m.dqhdt_p + dt3 * m.dqhdt_pp
    filt = -np.real(np.tensordot(m.Hi, np.conj(m.ph) * (m.filtr * rhs), axes=(0, 0))) \
        / m.H / m.dt / m.M ** 2
    unf = -np.real(np.tensordot(m.Hi, np.conj(m.ph) * rhs, axes=(0, 0))) / m.H / m.dt / m.M ** 2
    np.testing.assert_allclose(evaluate(m, "Dissspec"), filt - unf, atol=1e-20)


# ------------------------------------------------------------------ closures
PARAM_MODELS = {
    "qg": (OracleQG, dict(nx=32, dt=7200.0, **QG_PHYS), 21),
    "layered3": (OracleLayered, dict(nx=32, **LAYERED3), 22),
    "bt": (OracleBT, dict(nx=32, beta=1.5e-11, rd=15000.0, H=1.0, U=0.01, dt=3600.0), 23),
    "sqg": (OracleSQG, dict(nx=32, beta=0.0, H=1.0, U=0.0, dt=3600.0), 24),
}


def closure_state(name):
    """Oracle model of one golden closure case after three steps and an inversion."""
    from oracle.qg_diagnostics import attach_constants
    cls, kw, seed = PARAM_MODELS[name]
    m = cls(**kw)
    m.q = 1e-6 * np.random.default_rng(seed).standard_normal((m.nz, m.ny, m.nx))
    for _ in range(3):
        m._step_forward()
    m._invert()
    return attach_constants(m)


@pytest.mark.parametrize("name", list(PARAM_MODELS))
def test_closures_match_reference(name):
    """Smagorinsky, BackscatterBiharmonic and ZannaBolton2020 of the oracle
    equal the reference's on the same states (tests/golden/parameterizations.npz)."""
    import oracle.qg_parameterizations as P
    g = golden("parameterizations.npz")
    m = closure_state(name)
    assert rel_l2(m.q, g[f"{name}_q"]) < 1e-13
    du, dv = P.smagorinsky(m)
    assert rel_l2(du, g[f"{name}_smag_du"]) < 1e-12 and rel_l2(dv, g[f"{name}_smag_dv"]) < 1e-12
    assert rel_l2(P.smagorinsky(m, 0.2, just_viscosity=True), g[f"{name}_smag_nu"]) < 1e-12
    assert rel_l2(P.backscatter_biharmonic(m), g[f"{name}_back_dq"]) < 1e-12
    du, dv = P.zanna_bolton_2020(m)
    assert rel_l2(du, g[f"{name}_zb_du"]) < 1e-12 and rel_l2(dv, g[f"{name}_zb_dv"]) < 1e-12


@pytest.mark.parametrize("run", ["smag", "back", "zb", "combo"])
def test_closure_runs_match_reference(run):
    """Ten QGModel steps with a built-in closure as the model's parameterization."""
    import oracle.qg_parameterizations as P
    g = golden("parameterizations.npz")
    kinds = {
        "smag": ("uv_parameterization", lambda m: P.smagorinsky(m, 0.15), 10),
        "back": ("q_parameterization", lambda m: P.backscatter_biharmonic(m), 10),
        "zb": ("uv_parameterization", lambda m: P.zanna_bolton_2020(m), 10),
        "combo": ("uv_parameterization",
                  lambda m: np.sum([np.array(P.smagorinsky(m)) * 0.5,
                                    np.array(P.zanna_bolton_2020(m)) * 0.75], axis=0), 6),
    }
    kind, fn, steps = kinds[run]
    m = OracleQG(nx=32, dt=7200.0, **{kind: fn}, **QG_PHYS)
    from oracle.qg_diagnostics import attach_constants
    attach_constants(m)
    m.q = g["run_q0"]
    for _ in range(steps):
        m._step_forward()
    assert rel_l2(m.q, g[f"run_{run}_q"]) < 1e-12


def test_closure_planted_mistake_is_caught():
    """Smagorinsky with Sxy missing its factor 1/2 differs from the reference."""
    g = golden("parameterizations.npz")
    m = closure_state("qg")
    ik, il = 1j * m.k, 1j * m.l
    uh, vh = m.fft(m.u), m.fft(m.v)
    Sxx, Syy = m.ifft(uh * ik), m.ifft(vh * il)
    Sxy = m.ifft(uh * il + vh * ik)  # planted: no 0.5
    nu = (0.2 * m.dx) ** 2 * np.sqrt(2 * (Sxx ** 2 + Syy ** 2 + 2 * Sxy ** 2))
    assert rel_l2(nu, g["qg_smag_nu"]) > 1e-3
