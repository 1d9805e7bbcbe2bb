"""The learned-closure pipeline on the GPU: coarse-graining invariants, the
subgrid forcing against the oracle's restatement, the device-resident closure
step against the host-callback path, and what a trained network gains over the
coarse model without it (offline skill on held-out pairs)."""
import numpy as np
import pytest

from conftest import QG_PHYS, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mods(gpu):
    import torch  # noqa: F401
    import pyqg_amd
    from pyqg_amd.closure import cnn, data, param, train
    return pyqg_amd, cnn, data, param, train


def _oracle_forcing(q_hi, n_lo, phys, G):
    """NumPy restatement: S = C(T_hi(q)) - T_lo(C(q)) with the oracle's tendency."""
    from oracle.qg_oracle import OracleQG
    n_hi = q_hi.shape[-1]
    kw = dict(L=phys["L"], H1=phys["H1"], delta=phys["delta"], U1=phys["U1"], U2=phys["U2"],
              beta=phys["beta"], rd=phys["rd"], rek=phys["rek"], filterfac=phys["filterfac"])
    hi = OracleQG(nx=n_hi, **kw)
    hi.q = q_hi
    hi._invert()
    hi._do_advection()

    def coarsen(fh):
        out = np.zeros((fh.shape[0], n_lo, n_lo // 2 + 1), complex)
        for l in range(n_lo):
            if l == n_lo // 2:
                continue
            lw = l if l < n_lo // 2 else l - n_lo
            out[:, l, : n_lo // 2] = fh[:, lw % n_hi, : n_lo // 2]
        return out * (n_lo / n_hi) ** 2 * G

    lo = OracleQG(nx=n_lo, **kw)
    lo.qh = coarsen(hi.qh)
    lo._invert()
    lo._do_advection()
    S = coarsen(hi.dqhdt) - lo.dqhdt
    return lo.q, lo.u, lo.v, np.fft.irfftn(S, axes=(-2, -1))


def test_coarsen_is_a_projection_and_commutes_with_inversion(mods):
    pq = mods[0]
    from pyqg_amd import _lib
    hi = pq.Ensemble.qg(np.full(2, 128), full_state=True)
    lo = pq.Ensemble.qg(np.full(2, 64), full_state=True)
    rng = np.random.default_rng(0)
    hi.set_q([rng.standard_normal((2, 128, 128)) for _ in range(2)])
    hb, lb = hi.buckets[0], lo.buckets[0]
    _lib.check(_lib.lib.qgk_coarsen(hb.h, lb.h, _lib.FIELD_ID["qh"], _lib.FIELD_ID["qh"], None))
    once = lb.get("qh")
    # sharp truncation twice (hi -> lo -> lo) changes nothing: idempotent
    lo2 = pq.Ensemble.qg(np.full(2, 64), full_state=True)
    _lib.check(_lib.lib.qgk_coarsen(lb.h, lo2.buckets[0].h, _lib.FIELD_ID["qh"],
                                    _lib.FIELD_ID["qh"], None))
    assert np.array_equal(lo2.buckets[0].get("qh"), once)
    # commutes with the (diagonal-in-k) inversion: C(ph_hi) == a_lo C(qh_hi)
    _lib.check(_lib.lib.qgk_invert(hb.h))
    _lib.check(_lib.lib.qgk_invert(lb.h))
    _lib.check(_lib.lib.qgk_coarsen(hb.h, lo2.buckets[0].h, _lib.FIELD_ID["ph"],
                                    _lib.FIELD_ID["qh"], None))
    assert rel_l2(lo2.buckets[0].get("qh"), lb.get("ph")) < 1e-13
    # Parseval on the retained modes, exactly: the coarse field's mean square
    # on the coarse grid equals the mean square, on the fine grid, of the fine
    # field restricted to the modes the coarse grid keeps (numpy's transforms;
    # a wrong (n_lo/n_hi)^2 scale fails it)
    q_lo = lb.get("q")
    qh_hi = hb.get("qh")
    keep = np.zeros((128, 65), bool)
    for l in range(64):
        if l != 32:
            keep[(l if l < 32 else l - 64) % 128, :32] = True
    r_hi = np.fft.irfftn(np.where(keep, qh_hi, 0.0), s=(128, 128), axes=(-2, -1))
    ms_lo = (q_lo ** 2).mean(axis=(-2, -1))
    ms_hi = (r_hi ** 2).mean(axis=(-2, -1))
    assert np.all(np.abs(ms_lo - ms_hi) <= 1e-12 * ms_hi), (ms_lo, ms_hi)


def test_coarsen_plane_waves(mods):
    """A plane wave the coarse grid resolves comes through with its amplitude
    and phase: the coarse field is the same cosine sampled at every second
    fine grid point (coarse index j is fine index 2j); one beyond the coarse
    grid's band, or on its Nyquist row, disappears."""
    pq = mods[0]
    from pyqg_amd import _lib
    n_hi, n_lo, L = 128, 64, 1e6
    x_hi = (np.arange(n_hi) + 0.5) * L / n_hi
    two_pi = 2 * np.pi / L

    def wave(x, i, j, amp, phase):
        X, Y = np.meshgrid(x, x)
        return amp * np.cos(two_pi * (i * X + j * Y) + phase)

    cases = [((3, 5), True), ((31, -7), True), ((-12, 30), True), ((40, 2), False),
             ((5, 32), False), ((0, 50), False)]
    hi = pq.Ensemble.qg(np.full(len(cases), n_hi), full_state=True)
    lo = pq.Ensemble.qg(np.full(len(cases), n_lo), full_state=True)
    q_hi = [np.stack([wave(x_hi, i, j, 2.5, 0.3), wave(x_hi, i, j, -1.0, 1.1)])
            for (i, j), _ in cases]
    hi.set_q(q_hi)
    _lib.check(_lib.lib.qgk_coarsen(hi.buckets[0].h, lo.buckets[0].h, _lib.FIELD_ID["qh"],
                                    _lib.FIELD_ID["qh"], None))
    q_lo = lo.buckets[0].get("q")
    for m, ((i, j), kept) in enumerate(cases):
        if kept:
            want = q_hi[m][:, ::2, ::2]
            assert np.abs(q_lo[m] - want).max() < 1e-12 * 2.5, (i, j)
        else:
            assert np.abs(q_lo[m]).max() < 1e-12, (i, j)


def test_subgrid_forcing_matches_oracle(mods):
    pq, _, data, _, _ = mods
    gen = data.PairGenerator(nmembers=2, n_hi=128, n_lo=64, seed=3)
    gen.spin_up(30)
    q_hi = gen.hi.get("q")
    x, y = gen.sample()
    x, y = x.cpu().numpy(), y.cpu().numpy()
    for i in range(2):
        q, u, v, S = _oracle_forcing(q_hi[i], 64, gen.phys, gen.G)
        assert rel_l2(x[i, ..., 0:2].transpose(2, 0, 1), q) < 1e-6   # FP32 copies
        assert rel_l2(x[i, ..., 2:4].transpose(2, 0, 1), u) < 1e-6
        assert rel_l2(x[i, ..., 4:6].transpose(2, 0, 1), v) < 1e-6
        assert rel_l2(y[i].transpose(2, 0, 1), S) < 1e-6


def test_device_closure_step_equals_host_callback(mods):
    """The device-resident closure step and pyqg_amd.QGModel with the same
    network as a host-callback q_parameterization give the same trajectory."""
    pq, cnn, _, param, _ = mods
    net = cnn.CNN(channels=(6, 32, 16, 2), seed=1)
    sc = param.Scales(in_scale=[1e-6, 1e-6, 1e-2, 1e-2, 1e-2, 1e-2], out_scale=[1e-12, 1e-12])
    p = param.CNNParameterization(net, sc)
    phys = dict(QG_PHYS)
    ce = param.ClosedEnsemble(1, 64, 3600., phys, param=p, seed=4)
    q0 = ce.ens.get("q", 0).copy()
    m = pq.QGModel(nx=64, dt=3600., log_level=0, q_parameterization=p, **phys)
    m.q = q0
    for _ in range(5):
        ce.step(1)
        m._step_forward()
    # the network sees BF16 inputs in both paths; FP32 packing order differs
    assert rel_l2(ce.ens.get("q", 0), m.q) < 1e-4
    assert ce.ens.t == m.t


def test_closed_ensemble_graph_matches_step_by_step(mods):
    """ClosedEnsemble.step replays a graph of two steps (network on the
    bucket's stream into persistent buffers); in uneven chunks it matches the
    same ensemble stepped one step at a time bit for bit."""
    pq, cnn, _, param, _ = mods
    net = cnn.CNN(channels=(6, 32, 16, 2), seed=2)
    sc = param.Scales(in_scale=[1e-6, 1e-6, 1e-2, 1e-2, 1e-2, 1e-2], out_scale=[1e-12, 1e-12])
    p = param.CNNParameterization(net, sc)
    phys = dict(QG_PHYS)
    a = param.ClosedEnsemble(3, 64, 3600., phys, param=p, seed=5)
    b = param.ClosedEnsemble(3, 64, 3600., phys, param=p, seed=5)
    for chunk in (7, 6, 5):
        a.step(chunk)
        for _ in range(chunk):
            b._cnn_step()
        assert a.ens.tc == b.ens.tc and a.ens.t == b.ens.t
        for i in range(3):
            np.testing.assert_array_equal(a.ens.get("q", i), b.ens.get("q", i))


def _small_closure(cnn, param, seed):
    net = cnn.CNN(channels=(6, 32, 16, 2), seed=seed)
    sc = param.Scales(in_scale=[1e-6, 1e-6, 1e-2, 1e-2, 1e-2, 1e-2], out_scale=[1e-12, 1e-12])
    return param.CNNParameterization(net, sc)


def test_fused_closed_step_input_matches_packing(mods):
    """The fused row kernel writes the network input that qgk_pack_inputs
    forms from the per-phase real fields: the same BF16 values, but for the
    rare value whose rounding the transforms' last-bit differences move by one
    BF16 step."""
    import torch
    from pyqg_amd import _lib
    pq, cnn, _, param, _ = mods
    p = _small_closure(cnn, param, 3)
    a = param.ClosedEnsemble(4, 64, 3600., dict(QG_PHYS), param=p, seed=6, fused=True)
    b = param.ClosedEnsemble(4, 64, 3600., dict(QG_PHYS), param=p, seed=6, fused=False)
    assert a.fused and not b.fused
    L = _lib.lib
    _lib.check(L.qgk_fused_row_net(a.b.h, _lib.ptr(a.inv_in), a.cinp, param.ctypes_p(a.xin)))
    _lib.check(L.qgk_invert(b.b.h))
    _lib.check(L.qgk_pack_inputs(b.b.h, _lib.ptr(b.inv_in), b.cinp, param.ctypes_p(b.xin)))
    torch.cuda.synchronize()
    xa = a.xin.float().cpu().numpy()
    xb = b.xin.float().cpu().numpy()
    assert np.abs(xb[..., :6]).max() > 0
    assert np.all(xa[..., 6:] == 0) and np.all(xb[..., 6:] == 0)
    same = np.mean(xa == xb)
    assert same > 0.999, same
    step = np.abs(xb) * 2.0 ** -7 + 1e-30  # one BF16 step
    assert np.all(np.abs(xa - xb) <= step)


def test_fused_closed_step_matches_phases(mods):
    """The fused closed step (network input from the row kernel, forcing added
    in the column kernel) against the reference's phases one by one with the
    same network: the same trajectory up to rounding."""
    pq, cnn, _, param, _ = mods
    p = _small_closure(cnn, param, 4)
    a = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=p, seed=7, fused=True)
    b = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=p, seed=7, fused=False)
    for _ in range(6):
        a._cnn_step()
        b._cnn_step()
    for i in range(3):
        assert rel_l2(a.ens.get("q", i), b.ens.get("q", i)) < 1e-6
        assert rel_l2(a.ens.get("dqhdt_p", i), b.ens.get("dqhdt_p", i)) < 1e-6


def test_averaged_networks_closure(mods):
    """Several networks' forcings averaged in the closed step: two copies of
    one network give that network's trajectory; two different networks give,
    over one step from the same state, the mean of their two tendencies; the
    graph-replayed steps equal the step-by-step ones; the stochastic forcing
    is refused (each forcing add would draw it again)."""
    pq, cnn, _, param, _ = mods
    p1 = _small_closure(cnn, param, 4)
    p2 = param.CNNParameterization(_small_closure(cnn, param, 9).net, p1.scales)
    twice = param.CNNAverageParameterization([p1.net, p1.net], p1.scales)
    a = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=p1, seed=7, fused=False)
    b = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=twice, seed=7)
    assert not b.fused
    for _ in range(6):
        a._cnn_step()
        b._cnn_step()
    for i in range(3):
        assert rel_l2(a.ens.get("q", i), b.ens.get("q", i)) < 1e-12
    avg = param.CNNAverageParameterization([p1.net, p2.net], p1.scales)
    e1, e2, ea = (param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=p, seed=8,
                                       fused=False) for p in (p1, p2, avg))
    for e in (e1, e2, ea):
        e._cnn_step()
    for i in range(3):
        mean = 0.5 * (e1.ens.get("dqhdt_p", i) + e2.ens.get("dqhdt_p", i))
        assert rel_l2(ea.ens.get("dqhdt_p", i), mean) < 1e-10
        assert rel_l2(e1.ens.get("dqhdt_p", i), e2.ens.get("dqhdt_p", i)) > 1e-6
    g = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=avg, seed=8)
    h = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=avg, seed=8)
    g.step(10)
    for _ in range(10):
        h._cnn_step()
    for i in range(3):
        assert np.array_equal(g.ens.get("q", i), h.ens.get("q", i))
    with pytest.raises(ValueError):
        g.set_noise(np.zeros((2, 64, 33)), 0.5)


def test_fused_closed_step_without_forcing_is_the_plain_step(mods):
    """With the network's last layer zeroed the forcing is exactly zero, and
    the fused closed step is the plain fused step."""
    pq, cnn, _, param, _ = mods
    p = _small_closure(cnn, param, 5)
    p.net.w32[-1].zero_()
    p.net.b32[-1].zero_()
    p.net.refresh_bf16()
    a = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=p, seed=8, fused=True)
    b = param.ClosedEnsemble(3, 64, 3600., dict(QG_PHYS), param=None, seed=8)
    a.step(10)
    b.step(10)
    for i in range(3):
        assert rel_l2(a.ens.get("q", i), b.ens.get("q", i)) < 1e-12


def test_trained_network_beats_no_closure(mods):
    """Offline: on held-out pairs the trained network's forcing error is well
    below the no-closure prediction (zero forcing), i.e. R^2 > 0.3."""
    pq, cnn, data, param, train = mods
    gen = data.PairGenerator(nmembers=32, n_hi=128, n_lo=64, seed=0)
    gen.spin_up(24 * 300)
    x, y = gen.generate(1024, every=12)
    xt, yt = gen.generate(256, every=12)
    sc = param.Scales.fit(x, y)
    net = cnn.CNN(channels=(6, 64, 32, 32, 2), seed=0)
    before = train.evaluate(net, xt, yt, sc)
    info = train.train(net, x, y, sc, epochs=6, batch=32, lr=2e-3)
    after = train.evaluate(net, xt, yt, sc)
    print(f"offline R^2 untrained {before['r2']:.3f} -> trained {after['r2']:.3f} "
          f"({info['steps']} steps, {info['seconds']:.1f} s)", flush=True)
    assert after["r2"] > 0.3
    assert after["mse_net"] < 0.7 * after["mse_zero"]


def test_online_closure_at_the_experiments_resolution(mods):
    """The learned-closure pipeline at its real resolution, reduced in members
    and duration: 64 two-layer members at nx=256 (dt = 3600 s) spun up for five
    years and sampled daily (12 800 pairs of coarse inputs and subgrid forcing
    at nx=64), the CNN trained on them with the project's kernels, then 32
    coarse members at nx=64 (dt = 4 h) run for 720 days with and without the
    network.  Online, the network's run is closer to the nx=256 run than the
    plain coarse run: the time-mean KE spectrum of each layer (calc_ispec over
    the common wavenumbers) by a clear margin, and the energy-budget terms in
    pyqg's own comparison (diagnostic_similarities) on average."""
    from pyqg_amd.closure import experiment
    print("online closure: 64 x nx=256 -> nx=64", flush=True)
    r = experiment.run(members=64, samples=12800, spinup_days=1825, every=24, epochs=8,
                       online_members=32, online_days=720, avg_days=360, seed=1, closures=(),
                       log=lambda s: None)
    on = r["online"]
    assert r["offline"]["r2"] > 0.7
    for layer in ("KEspec1", "KEspec2"):
        assert on["cnn"]["ke_rel_err"][layer] < 0.8 * on["plain"]["ke_rel_err"][layer], layer
    assert on["cnn"]["budget_similarity_mean"] > 0.0
    assert on["cnn"]["finite"] and on["plain"]["finite"]


def _noise_draws(pq, amp, phi, ndraw, seed=11, M=8, n=64):
    """dqh after each of ndraw forcing calls with a zero network output: the
    stochastic term alone (qgk_set_spectral_noise), one draw per call."""
    import ctypes
    import torch
    from pyqg_amd import _lib
    ens = pq.Ensemble.qg(np.full(M, n), full_state=True, q_param=True)
    b = ens.buckets[0]
    b.set_spectral_noise(amp, phi, seed)
    net0 = torch.zeros((M, n, n, 4), dtype=torch.float32, device="cuda")
    sc = np.ones(2)
    out = []
    for _ in range(ndraw):
        _lib.check(_lib.lib.qgk_add_forcing_device(b.h, ctypes.c_void_p(net0.data_ptr()), 4,
                                                   _lib.ptr(sc)))
        out.append(b.get("dqh"))
    return np.array(out)             # (ndraw, M, nz, ny, nk)


def test_spectral_noise_variance_symmetry_and_fresh_draws(mods):
    """phi = 0: every point's draw has mean square amp^2 (here a spectrum that
    varies by layer and wavenumber), the kx = 0 and kx = nx/2 columns are the
    half spectrum of a real field (row ny - l the conjugate of row l, rows 0
    and ny/2 real), and successive calls draw afresh."""
    pq = mods[0]
    n = 64
    nk = n // 2 + 1
    l = np.fft.fftfreq(n, 1.0 / n)[:, None]
    k = np.arange(nk)[None, :]
    amp = np.stack([np.exp(-(k ** 2 + l ** 2) / 200.0) + 0.5, np.full((n, nk), 2.0)])
    d = _noise_draws(pq, amp, 0.0, 12)
    ms = (np.abs(d) ** 2).mean(axis=(0, 1))               # (nz, ny, nk)
    ratio = ms.sum(axis=(1, 2)) / (amp ** 2).sum(axis=(1, 2))
    np.testing.assert_allclose(ratio, 1.0, atol=0.02)
    # per wavenumber band, within the sampling error of 12 x 8 draws per point
    band = ms[:, 1:n // 2, 1:nk - 1].mean(axis=1) / (amp[:, 1:n // 2, 1:nk - 1] ** 2).mean(axis=1)
    assert np.abs(band - 1.0).max() < 0.12
    for kc in (0, n // 2):
        col = d[..., kc]                                   # (ndraw, M, nz, ny)
        np.testing.assert_array_equal(col[..., n - np.arange(1, n)], np.conj(col[..., 1:]))
        assert not col[..., 0].imag.any() and not col[..., n // 2].imag.any()
        assert np.abs(col[..., 0].real).max() > 0
    c = np.vdot(d[1:], d[:-1]).real / np.vdot(d, d).real
    assert abs(c) < 0.01


def test_spectral_noise_ar1_correlation_and_seed(mods):
    """phi = 0.8: after the start-up the process keeps the prescribed mean
    square, and one call to the next correlates by phi; the same seed draws
    the same sequence, another seed another."""
    pq = mods[0]
    n, nk = 64, 33
    amp = np.full((2, n, nk), 3.0)
    d = _noise_draws(pq, amp, 0.8, 50)
    s = d[20:]
    ms = (np.abs(s) ** 2).mean()
    assert abs(ms / 9.0 - 1.0) < 0.03
    rho = np.vdot(s[:-1], s[1:]).real / np.vdot(s[:-1], s[:-1]).real
    assert abs(rho - 0.8) < 0.02
    again = _noise_draws(pq, amp, 0.8, 3)
    np.testing.assert_array_equal(again, d[:3])
    other = _noise_draws(pq, amp, 0.8, 3, seed=12)
    assert np.abs(other - d[:3]).max() > 1.0


def test_closed_ensemble_with_noise_graph_matches_step_by_step(mods):
    """With the stochastic forcing on, the replayed step graphs draw the same
    numbers as steps taken one at a time (the step counter lives on the
    device), and the forcing changes the trajectory."""
    pq, cnn, _, param, _ = mods
    p = _small_closure(cnn, param, 4)
    phys = dict(QG_PHYS)
    amp = np.full((2, 64, 33), 1e-9)
    a = param.ClosedEnsemble(3, 64, 3600., phys, param=p, seed=5)
    b = param.ClosedEnsemble(3, 64, 3600., phys, param=p, seed=5)
    c = param.ClosedEnsemble(3, 64, 3600., phys, param=p, seed=5)
    a.set_noise(amp, 0.5, seed=3)
    b.set_noise(amp, 0.5, seed=3)
    for chunk in (7, 6):
        a.step(chunk)
        c.step(chunk)
        for _ in range(chunk):
            b._cnn_step()
        for i in range(3):
            np.testing.assert_array_equal(a.ens.get("q", i), b.ens.get("q", i))
    assert np.abs(a.ens.get("q", 0) - c.ens.get("q", 0)).max() > 0
