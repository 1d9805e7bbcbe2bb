"""The drop-in PseudoSpectralKernel on the GPU: the reference's own closed-form
kernel tests (pyqg/tests/test_model.py, test_advection.py, test_fft2.py)
restated against pyqg_amd's models, the transforms against numpy.fft, and the
golden vectors of the reference build."""
import numpy as np
import pytest

from conftest import QG_PHYS, golden, rel_l2

pytestmark = pytest.mark.gpu

SIZES = [8, 16, 32, 64, 128, 256, 512, 1024]


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


def closed_form_model(pq, **kw):
    return pq.QGModel(beta=0., U1=0., U2=0., filterfac=0., log_level=0, **kw)


# --------------------------------------------------------------- transforms
@pytest.mark.parametrize("n", SIZES)
def test_fft_ifft_match_numpy(pq, n):
    print(f"fft n={n}", flush=True)
    k = pq.PseudoSpectralKernel(2, n, n)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((2, n, n))
    X = k.fft(x)
    ref = np.fft.rfftn(x, axes=(-2, -1))
    assert np.abs(X - ref).max() <= 1e-14 * np.abs(ref).max() * np.log2(n)
    # arbitrary (non-Hermitian) input: c2r must drop Im of the kx=0, nx/2 columns
    Y = rng.standard_normal((2, n, n // 2 + 1)) + 1j * rng.standard_normal((2, n, n // 2 + 1))
    y = k.ifft(Y)
    yref = np.fft.irfftn(Y, axes=(-2, -1))
    assert np.abs(y - yref).max() <= 1e-14 * np.abs(yref).max() * np.log2(n)
    # round trip
    np.testing.assert_allclose(k.ifft(k.fft(x)), x, rtol=0, atol=1e-13 * np.abs(x).max())


def test_fft_bin_amplitudes(pq):
    """test_model.py:16-65"""
    m = closed_form_model(pq)
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
            assert np.abs(a).max() <= 1e-15 * amp


def test_inversion_barotropic(pq):
    """test_model.py:67-94"""
    m = closed_form_model(pq)
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            q = np.cos(k * m.x) + np.sin(l * m.y)
            m.set_q1q2(q, q)
            m._invert()
            for z in range(m.nz):
                np.testing.assert_allclose(m.u[z], l ** -1 * np.cos(l * m.y), 1e-13)
                np.testing.assert_allclose(m.v[z], k ** -1 * np.sin(k * m.x), 1e-13)


def test_change_inversion_matrix(pq):
    """test_model.py:101-106"""
    m = closed_form_model(pq)
    a_new = np.random.rand(m.nz, m.nz, m.nl, m.nk)
    m.a = a_new
    np.testing.assert_allclose(a_new, m.a)


def test_advection_through_getter_views(pq):
    """test_model.py:108-158: u, v written through the getters' arrays."""
    m = closed_form_model(pq)
    amp = (m.nx / 2) ** 2 * m.ny ** 2
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            m.set_q1q2(np.cos(k * m.x), np.sin(l * m.y))
            m.u[0] = np.sin(k * m.x)
            m.v[0] = np.zeros_like(m.y)
            m.u[1] = np.zeros_like(m.x)
            m.v[1] = np.cos(l * m.y)
            m.set_U1U2(0., 0.)
            m._do_advection()
            t = np.real(m.dqhdt * m.dqhdt.conj())
            np.testing.assert_allclose(t[0, 0, 2 * kw], k ** 2 * amp, 1e-14)
            np.testing.assert_allclose(t[1, 2 * lw, 0], l ** 2 * amp, 1e-14)
            np.testing.assert_allclose(t[1, -2 * lw, 0], l ** 2 * amp, 1e-14)
            t[0, 0, 2 * kw] = t[1, 2 * lw, 0] = t[1, -2 * lw, 0] = 0
            assert np.abs(t).max() <= 1e-14


def test_friction(pq):
    """test_model.py:160-183"""
    m = closed_form_model(pq)
    for kw in range(1, m.nx // 8):
        for lw in range(1, m.ny // 8):
            k, l = 2 * np.pi * kw / m.L, 2 * np.pi * lw / m.W
            m.set_q1q2(np.cos(k * m.x), np.sin(l * m.y))
            m._invert()
            m.dqhdt[:] = 0.
            m._do_friction()
            np.testing.assert_allclose(m.dqhdt[-1], m.rek * m.wv2 * m.ph[-1], 1e-15)
            np.testing.assert_allclose(m.dqhdt[:-1], 0., 1e-15)


def test_q_parameterization_adds_one(pq):
    """test_model.py:186-205"""
    m1 = closed_form_model(pq)
    m2 = closed_form_model(pq, q_parameterization=lambda m: np.ones_like(m1.q))
    m2.q = m1.q
    m1._step_forward()
    m2._step_forward()
    np.testing.assert_allclose(m2.ifft(m2.dqhdt) - m1.ifft(m1.dqhdt), 1., 1e-15)


def test_uv_parameterization_zero_and_nonzero_curl(pq):
    """test_model.py:208-262"""
    m1 = closed_form_model(pq)
    const = lambda m: (np.ones_like(m.u) * 0.3 + 0.7, np.ones_like(m.v) * -0.2 + 0.1)
    m2 = closed_form_model(pq, uv_parameterization=const)
    m2.q = m1.q
    m1._step_forward()
    m2._step_forward()
    np.testing.assert_allclose(m2.ifft(m2.dqhdt) - m1.ifft(m1.dqhdt), 0., 1e-15)
    rng = np.random.default_rng(0)
    noisy = lambda m: (rng.normal(size=m.u.shape), rng.normal(size=m.v.shape))
    m3 = closed_form_model(pq, uv_parameterization=noisy)
    m3.q = m1.q
    m3._step_forward()
    assert np.abs(m3.ifft(m3.dqhdt) - m1.ifft(m1.dqhdt)).max() > 1e-10


def test_parameterization_variables(pq):
    """test_model.py:265-286: dqh / duh / dvh exist only when configured."""
    m1 = closed_form_model(pq)
    m2 = closed_form_model(pq, q_parameterization=lambda m: None)
    m3 = closed_form_model(pq, uv_parameterization=lambda m: None)
    m2.dqh
    m3.duh
    m3.dvh
    with pytest.raises(AttributeError):
        m1.dqh
    with pytest.raises(AttributeError):
        m1.duh
    with pytest.raises(AttributeError):
        m2.dvh


def test_parameterization_type_inference(pq):
    """test_model.py:289-311"""
    class Q:
        parameterization_type = "q_parameterization"

        def __call__(self, m):
            return np.zeros_like(m.q)

    q = Q()
    m = closed_form_model(pq, parameterization=q)
    assert m.q_parameterization is q
    with pytest.raises(ValueError):
        closed_form_model(pq, parameterization="boop")
    with pytest.raises(AssertionError):
        closed_form_model(pq, q_parameterization=q, parameterization=q)


def test_timestepping_startup(pq):
    """test_model.py:314-340 (tendency written through the getter)."""
    m = closed_form_model(pq)
    m.set_q(np.zeros_like(m.q))
    d = np.random.rand(*m.dqhdt.shape) + 1j * np.random.rand(*m.dqhdt.shape)
    m.dqhdt[:] = d
    assert m.tc == 0
    for n in (1, 2, 3):
        m._forward_timestep()
        np.testing.assert_allclose(m.qh, n * m.dt * d)


def test_typed_setters_raise_like_cython(pq):
    m = closed_form_model(pq)
    with pytest.raises(ValueError):
        m.Qy = np.array([1, 2])          # int64: buffer dtype mismatch
    with pytest.raises(ValueError):
        m.a = np.zeros((2, 2, m.nl))     # wrong ndim
    with pytest.raises(ValueError):
        m.fft(np.zeros((2, m.ny, m.nx), np.complex128))


def test_plane_wave_jacobian_vanishes(pq):
    """test_advection.py:4-52 at nx=256."""
    m = pq.QGModel(L=2. * np.pi, nx=256, U1=0, log_level=0)
    for kx, ky in [(1., 2.), (5., 0.), (10., 10.), (0., 21.), (24., 12.), (2., 49.)]:
        m.set_q1q2(np.cos(kx * m.x + ky * m.y), np.zeros_like(m.x))
        m._invert()
        jac = m.ifft(m._advect(m.q, m.u, m.v))
        assert np.abs(jac).sum() * m.dx * m.dy / m.L ** 2 < 1e-13


@pytest.mark.parametrize("cls", ["BTModel", "QGModel"])
def test_parseval_and_round_trip(pq, cls):
    """test_fft2.py:8-40"""
    if cls == "BTModel":
        m = pq.BTModel(L=2. * np.pi, beta=0., U=0., rd=0., filterfac=0, log_level=0)
    else:
        m = pq.QGModel(L=2. * np.pi, beta=0., U1=0., U2=0., filterfac=0., log_level=0)
    p = np.random.randn(m.nz, m.ny, m.nx)
    p -= p.mean(axis=(-2, -1))[:, np.newaxis, np.newaxis]
    ph = m.fft(p)
    v = 2. * np.abs(ph) ** 2 / m.M ** 2
    v[..., 0] /= 2
    v[..., -1] /= 2
    err = np.abs((p.var(axis=(-2, -1)) - v.sum(axis=(-2, -1))) / p.var(axis=(-2, -1))).sum()
    assert err < 1e-14
    assert np.allclose(p, m.ifft(ph), rtol=1e-5)


# ------------------------------------------------------------------ golden
STATE = ["q", "qh", "ph", "u", "v", "uq", "vq", "uqh", "vqh", "dqhdt", "dqhdt_p", "dqhdt_pp"]


def test_golden_startup_every_field(pq):
    g = golden("qg_nx32_startup.npz")
    m = pq.QGModel(nx=32, dt=7200.0, log_level=0, **QG_PHYS)
    for f in ["kk", "ll", "filtr", "Qy", "Ubg", "_k2l2"]:
        np.testing.assert_allclose(getattr(m, f), g["const_" + f], rtol=1e-15, atol=0)
    np.testing.assert_allclose(m.a.real, g["const_a"], rtol=1e-14, atol=0)
    m.q = g["q0"]
    assert rel_l2(m.qh, g["qh0"]) < 1e-14
    for step in range(1, 4):
        m._step_forward()
        if step in (1, 3):
            for f in STATE:
                ref = g[f"s{step}_{f}"]
                if np.abs(ref).max() == 0:
                    assert np.abs(getattr(m, f)).max() == 0
                    continue
                assert rel_l2(getattr(m, f), ref) < 1e-13, f
            assert m.t == g[f"s{step}_t"] and m.tc == g[f"s{step}_tc"]
            assert m.ablevel == g[f"s{step}_ablevel"]


def q_param(m):
    return -2e-7 * m.q + 1e-13 * np.cos(2 * np.pi * m.x / m.L)[np.newaxis]


def uv_param(m):
    return 1e-6 * m.v, -1e-6 * m.u + 1e-9 * np.sin(2 * np.pi * m.y / m.W)[np.newaxis]


def test_golden_parameterized(pq):
    g = golden("qg_param_nx32.npz")
    m = pq.QGModel(nx=32, dt=7200.0, log_level=0, q_parameterization=q_param,
                   uv_parameterization=uv_param, **QG_PHYS)
    m.q = g["q0"]
    for _ in range(3):
        m._step_forward()
    for f in ("q", "qh", "dqhdt", "dqh", "duh", "dvh"):
        assert rel_l2(getattr(m, f), g["s3_" + f]) < 1e-12, f


def test_golden_bt_sqg(pq):
    g = golden("bt_sqg_nx64.npz")
    m = pq.BTModel(L=2. * np.pi, nx=64, beta=0., H=1., rek=0., rd=None, dt=0.0025,
                   log_level=0)
    m.q = g["bt_q0"]
    for _ in range(100):
        m._step_forward()
    assert rel_l2(m.q, g["bt_s100_q"]) < 1e-11
    assert m.t == g["bt_s100_t"]
    m = pq.SQGModel(L=2. * np.pi, nx=64, beta=0., H=1., dt=0.0025, log_level=0)
    m.q = g["sqg_q0"]
    m._fused_steps(100)
    assert rel_l2(m.q, g["sqg_s100_q"]) < 1e-11


def test_golden_layered(pq):
    g = golden("layered_nz3_nx64.npz")
    m = pq.LayeredModel(nx=64, log_level=0, nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0],
                        H=[500, 1750, 1750], rho=[1025.0, 1025.3, 1025.6], L=1e6,
                        rek=5.787e-7, g=9.81, dt=3600.0, beta=1.5e-11, rd=15000.0,
                        f=0.0001236812857687059)
    np.testing.assert_allclose(m.S, g["const_S"], rtol=1e-15)
    np.testing.assert_allclose(m.a[:, :, 5, 5].real, g["const_a_5_5"], rtol=1e-12)
    m.q = g["q0"]
    m._step_forward()
    for f in ("q", "ph", "dqhdt"):
        assert rel_l2(getattr(m, f), g["s1_" + f]) < 1e-12, f
    for _ in range(19):
        m._step_forward()
    assert rel_l2(m.q, g["s20_q"]) < 1e-11


def test_golden_layered_nz10(pq):
    g = golden("layered_nz10_nx32.npz")
    m = pq.LayeredModel(nz=10, nx=32, H=g["in_H"], rho=g["in_rho"], U=g["in_U"],
                        V=np.zeros(10), L=2e6, beta=1.5e-11, rek=5.787e-7, dt=3600.0,
                        f=0.0001236812857687059, g=9.81, log_level=0)
    np.testing.assert_allclose(m.radii, g["const_radii"], rtol=1e-10)
    np.testing.assert_allclose(m.a[:, :, 5, 5].real, g["const_a_5_5"], rtol=1e-10, atol=0)
    m.q = g["q0"]
    m._fused_steps(10)
    assert rel_l2(m.q, g["s10_q"]) < 1e-10


def test_bt_long_run_reference_values(pq):
    """test_reference_solns.py:97-135: 2000 steps (fused path), exact time."""
    lr = golden("long_runs.npz")
    m = pq.BTModel(L=2. * np.pi, nx=64, tmax=5, beta=0., H=1., rek=0., rd=None, dt=0.0025,
                   twrite=1000, log_level=0)
    p = (np.exp(-(2. * (m.x - 1.75 * np.pi / 2)) ** 2. - (2. * (m.y - np.pi)) ** 2)
         + np.exp(-(2. * (m.x - 2.25 * np.pi / 2)) ** 2. - (2. * (m.y - np.pi)) ** 2))
    ph = m.fft(p[np.newaxis])
    KEaux = m.spec_var(m.filtr * m.wv * ph) / 2.
    qi = m.ifft(-m.wv2 * (ph / np.sqrt(KEaux)))
    m.set_q(qi)
    np.testing.assert_allclose(m.q, qi, 1e-5, 1e-14)
    m.run()
    assert m.t == 5.000000000000082
    np.testing.assert_allclose((m.q ** 2).sum(), 89101.741238768518, 1e-5)
    np.testing.assert_allclose((m.ifft(m.ph) ** 2).sum(), 1493.217664248918, 1e-5)
    np.testing.assert_allclose(m._calc_ke(), 0.9950360837282386, 1e-5)
    np.testing.assert_allclose((m.q ** 2).sum(), lr["bt_qnorm"], rtol=1e-9)


# --------------------------------------------------------------- runtime
@pytest.mark.parametrize("order", ["kernels_first", "torch_first"])
def test_one_hip_runtime_with_torch(gpu, order):
    """libqgk and torch share one HIP runtime whichever is imported first
    (two runtimes in one process leave the second without a device)."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    body = ("import numpy as np, pyqg_amd\n"
            "e = pyqg_amd.Ensemble.qg(np.full(2, 64), dt=3600.)\n"
            "e.set_q([np.zeros((2, 64, 64))] * 2); e.step(3)\n")
    torch_op = ("import torch\n"
                "x = torch.ones(4, device='cuda'); torch.cuda.synchronize()\n"
                "assert float(x.sum()) == 4.0\n")
    code = body + torch_op if order == "kernels_first" else torch_op + body + torch_op
    code += ("maps = {l.split()[-1] for l in open('/proc/self/maps') if 'amdhip64' in l}\n"
             "assert len(maps) == 1, maps\nprint('ok')\n")
    print(f"runtime order {order}", flush=True)
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=180)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stderr[-2000:]


def test_run_with_snapshots_matches_oracle(pq):
    """model.py:260-278: yields after the steps with t >= tsnapstart and
    tc % ceil(tsnapint / dt) == 0; the yielded states are the per-step path's."""
    from oracle.qg_oracle import OracleQG
    dt = 3600.0
    kw = dict(nx=64, dt=dt, **QG_PHYS)
    q0 = 1e-7 * np.random.default_rng(8).standard_normal((2, 64, 64))
    m = pq.QGModel(log_level=0, tmax=30 * dt, **kw)
    m.q = q0
    o = OracleQG(**kw)
    o.q = q0
    got = []
    for t in m.run_with_snapshots(tsnapstart=10 * dt, tsnapint=7 * dt):
        got.append((t, m.tc, m.q.copy(), m.ph.copy()))
    want = []
    while o.t < 30 * dt:
        o._step_forward()
        if o.t >= 10 * dt and o.tc % 7 == 0:
            want.append((o.t, o.tc, o.q.copy(), o.ph.copy()))
    assert [(t, tc) for t, tc, _, _ in got] == [(t, tc) for t, tc, _, _ in want]
    assert len(got) == 3
    for (_, _, q, ph), (_, _, qo, pho) in zip(got, want):
        assert rel_l2(q, qo) < 1e-12 and rel_l2(ph, pho) < 1e-12
    assert m.t == o.t and m.tc == o.tc


# ------------------------------------------------------- rectangular grids
def _rect_constants(nz, ny, nx, L=1e6):
    """Wavenumbers, inversion and background of an nz-layer channel on an
    ny x nx grid of square cells.  The model classes build ll from nx
    (model.py:461-462) and are square-only; the kernel sizes every array from
    nz, ny and nx (kernel.pyx:120-206), so a caller with its own constants can
    step rectangular grids.  a = (S - kappa^2 I)^-1 per wavenumber, zero at
    kappa = 0, for the stretching S of nz layers with interface couplings F."""
    W = L * ny / nx
    kk = 2 * np.pi / L * np.arange(nx // 2 + 1)
    ll = 2 * np.pi / W * np.fft.fftfreq(ny, 1.0 / ny)
    k2l2 = kk[np.newaxis, :] ** 2 + ll[:, np.newaxis] ** 2
    S = np.zeros((nz, nz))
    if nz == 1:
        S[0, 0] = -(1 / 15e3) ** 2
    for i, F in enumerate([4.4e-9, 2.1e-9][:nz - 1]):
        S[i, i] -= F
        S[i, i + 1] += F
        S[i + 1, i] += 0.6 * F
        S[i + 1, i + 1] -= 0.6 * F
    M = S[:, :, np.newaxis, np.newaxis] - k2l2 * np.eye(nz)[:, :, np.newaxis, np.newaxis]
    M[:, :, k2l2 == 0] = np.eye(nz)[:, :, np.newaxis]   # S is singular at kappa = 0
    a = np.linalg.inv(M.transpose(2, 3, 0, 1)).transpose(2, 3, 0, 1)
    a[:, :, k2l2 == 0] = 0.0
    dx = L / nx
    wvx = np.sqrt(k2l2) * dx
    filtr = np.where(wvx <= 0.65 * np.pi, 1.0, np.exp(-23.6 * (wvx - 0.65 * np.pi) ** 4))
    Ubg = np.linspace(0.025, 0.0, nz)
    Qy = 1.5e-11 - S @ Ubg
    return dict(kk=kk, ll=ll, a=a, filtr=filtr, Ubg=Ubg, Qy=Qy)


@pytest.mark.parametrize("nz,ny,nx", [(2, 64, 128), (2, 128, 32), (2, 96, 64), (1, 32, 48),
                                      (3, 64, 32)])
def test_rectangular_kernel_matches_oracle_kernel(pq, nz, ny, nx):
    """The bare kernel on an ny != nx grid against the oracle's kernel with the
    same constants: 10 steps method by method (every intermediate field), then
    30 fused steps (the row transforms at length nx, the column transforms at
    length ny)."""
    from oracle.qg_oracle import OracleKernel
    print(f"rectangular kernel nz={nz} ny={ny} nx={nx}", flush=True)
    k = pq.PseudoSpectralKernel(nz, ny, nx)
    o = OracleKernel(nz, ny, nx)
    c = _rect_constants(nz, ny, nx)
    q0 = 1e-6 * np.random.default_rng(ny + nx + nz).standard_normal((nz, ny, nx))
    for m in (k, o):
        m.kk = c["kk"]
        m.ll = c["ll"]
        m.a = c["a"]
        m.filtr = c["filtr"]
        m.Ubg = c["Ubg"]
        m.Qy = c["Qy"]
        m.rek = 5.787e-7
        m.dt = 3600.0
        m.q = q0
    np.testing.assert_array_equal(k._k2l2, o._k2l2)
    for _ in range(10):
        for m in (k, o):
            m._invert()
            m._do_advection()
            m._do_friction()
            m._forward_timestep()
    for f in ("q", "qh", "ph", "u", "v", "uq", "vq", "uqh", "vqh", "dqhdt", "dqhdt_p"):
        assert rel_l2(getattr(k, f), getattr(o, "_" + f)) < 1e-12, f
    k._fused_steps(30)
    for _ in range(30):
        o._invert()
        o._do_advection()
        o._do_friction()
        o._forward_timestep()
    assert k.tc == o.tc == 40 and k.t == o.t
    assert rel_l2(k.q, o._q) < 1e-11
    assert rel_l2(k.qh, o._qh) < 1e-11
    assert rel_l2(k.ph, o._ph) < 1e-11
