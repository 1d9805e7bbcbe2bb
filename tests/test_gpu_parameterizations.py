"""Built-in closures on the device (pyqg/parameterizations.py:132-360) against
the reference's outputs on the same states and runs
(tests/golden/parameterizations.npz), plus the reference's own checks
(pyqg/tests/test_parameterizations.py)."""
import numpy as np
import pytest

from conftest import QG_PHYS, golden, rel_l2

pytestmark = pytest.mark.gpu

LAYERED3 = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
                rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
                beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)
MODELS = {
    "qg": ("QGModel", dict(nx=32, dt=7200.0, **QG_PHYS), 21),
    "layered3": ("LayeredModel", dict(nx=32, **LAYERED3), 22),
    "bt": ("BTModel", dict(nx=32, beta=1.5e-11, rd=15000.0, H=1.0, U=0.01, dt=3600.0), 23),
    "sqg": ("SQGModel", dict(nx=32, beta=0.0, H=1.0, U=0.0, dt=3600.0), 24),
}


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


def state(pq, name):
    cls, kw, seed = MODELS[name]
    m = getattr(pq, cls)(log_level=0, **kw)
    m.q = 1e-6 * np.random.default_rng(seed).standard_normal((m.nz, m.ny, m.nx))
    for _ in range(3):
        m._step_forward()
    m._invert()
    return m


@pytest.mark.parametrize("name", list(MODELS))
def test_closures_match_reference(pq, name):
    print(f"closures {name}", flush=True)
    g = golden("parameterizations.npz")
    m = state(pq, name)
    assert rel_l2(m.q, g[f"{name}_q"]) < 1e-12
    du, dv = pq.Smagorinsky()(m)
    assert rel_l2(du, g[f"{name}_smag_du"]) < 1e-11 and rel_l2(dv, g[f"{name}_smag_dv"]) < 1e-11
    assert rel_l2(pq.Smagorinsky(0.2)(m, just_viscosity=True), g[f"{name}_smag_nu"]) < 1e-11
    assert rel_l2(pq.BackscatterBiharmonic()(m), g[f"{name}_back_dq"]) < 1e-11
    du, dv = pq.ZannaBolton2020()(m)
    assert rel_l2(du, g[f"{name}_zb_du"]) < 1e-11 and rel_l2(dv, g[f"{name}_zb_dv"]) < 1e-11


@pytest.mark.parametrize("run", ["smag", "back", "zb", "combo"])
def test_runs_with_device_closures_match_reference(pq, run):
    """The closure as the model's parameterization: forcing written into the
    kernel's du/dv or dq on the device, no host round trip."""
    print(f"closure run {run}", flush=True)
    g = golden("parameterizations.npz")
    make = {
        "smag": ("uv_parameterization", lambda: pq.Smagorinsky(0.15), 10),
        "back": ("q_parameterization", lambda: pq.BackscatterBiharmonic(), 10),
        "zb": ("uv_parameterization", lambda: pq.ZannaBolton2020(), 10),
        "combo": ("uv_parameterization",
                  lambda: 0.5 * pq.Smagorinsky() + 0.75 * pq.ZannaBolton2020(), 6),
    }
    kind, fn, steps = make[run]
    m = pq.QGModel(nx=32, dt=7200.0, log_level=0, **{kind: fn()}, **QG_PHYS)
    m.q = g["run_q0"]
    for _ in range(steps):
        m._step_forward()
    assert rel_l2(m.q, g[f"run_{run}_q"]) < 1e-11


def test_device_path_equals_host_callback(pq):
    """A closure run in place on the device and the same closure returned to
    the host through the callback path of kernel.pyx:397-422 step alike."""
    q0 = 1e-6 * np.random.default_rng(5).standard_normal((2, 32, 32))
    smag = pq.Smagorinsky(0.12)
    a = pq.QGModel(nx=32, log_level=0, uv_parameterization=smag)
    b = pq.QGModel(nx=32, log_level=0, uv_parameterization=lambda m: smag(m))
    for m in (a, b):
        m.q = q0
        for _ in range(5):
            m._step_forward()
    assert rel_l2(a.q, b.q) < 1e-14


def test_addition_and_scaling(pq):
    """pyqg/tests/test_parameterizations.py:39-52"""
    for name in MODELS:
        m = state(pq, name)
        back = pq.BackscatterBiharmonic()
        smag = pq.Smagorinsky()
        zb20 = pq.ZannaBolton2020()
        comb = 0.5 * smag + 0.75 * zb20
        du, dv = comb(m)
        np.testing.assert_allclose(du, 0.5 * smag(m)[0] + 0.75 * zb20(m)[0], rtol=1e-11)
        np.testing.assert_allclose(dv, 0.5 * smag(m)[1] + 0.75 * zb20(m)[1], rtol=1e-11)
        with pytest.raises(AssertionError):
            back + smag


def test_ring_forcing(pq):
    """RingForcing (parameterizations.py:291-360): zero-mean, unit mean
    absolute value per layer times the magnitude."""
    m = state(pq, "qg")
    ring = pq.RingForcing(k_in_forc=3 * m.dk, k_out_forc=6 * m.dk, mag_noise_forc=2.5)
    dq = ring(m)
    assert dq.shape == (2, 32, 32)
    np.testing.assert_allclose(dq.mean(axis=(-1, -2)), 0.0, atol=1e-12)
    np.testing.assert_allclose(np.abs(dq).mean(axis=(-1, -2)), 2.5, rtol=1e-12)
    m2 = pq.QGModel(nx=32, log_level=0, q_parameterization=ring * 1e-12)
    m2.q = m.q
    for _ in range(3):
        m2._step_forward()
    assert np.isfinite(m2.q).all()


@pytest.mark.parametrize("which", ["smag", "back"])
def test_ensemble_closure_matches_single_models(pq, which):
    """Ensemble.step_parameterized runs the device closure for every member as
    each member's own QGModel with that parameterization does."""
    M, nx, steps = 3, 32, 6
    make = {"smag": lambda: pq.Smagorinsky(0.15), "back": lambda: pq.BackscatterBiharmonic()}
    p = make[which]()
    uv = p.parameterization_type == "uv_parameterization"
    betas = np.array([1.0e-11, 1.5e-11, 2.0e-11])
    ens = pq.Ensemble.qg(nx, beta=betas, dt=3600.0, nmembers=M, full_state=True,
                         q_param=not uv, uv_param=uv)
    q0 = [1e-6 * np.random.default_rng(40 + i).standard_normal((2, nx, nx)) for i in range(M)]
    ens.set_q(q0)
    ens.step_parameterized(steps, p)
    phys = {k: v for k, v in QG_PHYS.items() if k != "beta"}
    for i in range(M):
        kind = "uv_parameterization" if uv else "q_parameterization"
        m = pq.QGModel(nx=nx, dt=3600.0, beta=betas[i], log_level=0, **{kind: make[which]()},
                       **phys)
        m.q = q0[i]
        for _ in range(steps):
            m._step_forward()
        assert rel_l2(ens.get("q", i), m.q) < 1e-13, i


def test_closures_on_a_larger_grid_match_oracle(pq):
    """nx=128, past the one-launch plane transforms: the closures' derivative
    fields are formed by their own pointwise kernels and transformed by row
    and column launches that carry several fields.  Closure outputs on one
    state, and 8 steps with each closure as the model's parameterization,
    against the oracle."""
    from oracle.qg_diagnostics import attach_constants
    from oracle.qg_oracle import OracleQG
    import oracle.qg_parameterizations as P
    print("closures nx=128", flush=True)
    kw = dict(nx=128, dt=3600.0, **QG_PHYS)
    q0 = 1e-6 * np.random.default_rng(128).standard_normal((2, 128, 128))
    m = pq.QGModel(log_level=0, **kw)
    m.q = q0
    m._invert()
    o = OracleQG(**kw)
    o.q = q0
    o._invert()
    attach_constants(o)
    du, dv = pq.Smagorinsky()(m)
    odu, odv = P.smagorinsky(o)
    assert rel_l2(du, odu) < 1e-11 and rel_l2(dv, odv) < 1e-11
    assert rel_l2(pq.Smagorinsky(0.2)(m, just_viscosity=True),
                  P.smagorinsky(o, 0.2, just_viscosity=True)) < 1e-11
    assert rel_l2(pq.BackscatterBiharmonic()(m), P.backscatter_biharmonic(o)) < 1e-11
    du, dv = pq.ZannaBolton2020()(m)
    odu, odv = P.zanna_bolton_2020(o)
    assert rel_l2(du, odu) < 1e-11 and rel_l2(dv, odv) < 1e-11
    runs = {
        "smag": ("uv_parameterization", lambda: pq.Smagorinsky(0.15),
                 lambda mm: P.smagorinsky(mm, 0.15)),
        "back": ("q_parameterization", lambda: pq.BackscatterBiharmonic(),
                 lambda mm: P.backscatter_biharmonic(mm)),
        "zb": ("uv_parameterization", lambda: pq.ZannaBolton2020(),
               lambda mm: P.zanna_bolton_2020(mm)),
    }
    for name, (kind, dev, ref) in runs.items():
        a = pq.QGModel(log_level=0, **{kind: dev()}, **kw)
        b = OracleQG(**{kind: ref}, **kw)
        attach_constants(b)
        a.q = q0
        b.q = q0
        for _ in range(8):
            a._step_forward()
            b._step_forward()
        assert rel_l2(a.q, b.q) < 1e-10, name
