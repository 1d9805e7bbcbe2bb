"""Device-accumulated diagnostics against the reference and the oracle.

* The time means of every registered diagnostic after a few steps of each
  model class equal those written from the reference build
  (tests/golden/diagnostics.npz; pyqg/model.py:644-822, qg_diagnostics.py,
  qg_model.py:213-263, layered_model.py:242-301).
* Model.run() with fused stretches between accumulation steps gives the
  oracle's means; Ensemble.accumulate_diagnostics gives every member's.
* The reference's own identities (pyqg/tests/test_diagnostics.py:29-160).
"""
import numpy as np
import pytest

from conftest import QG_PHYS, golden, rel_l2

pytestmark = pytest.mark.gpu

LAYERED3 = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
                rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
                beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)

CASES = {
    "qg": ("QGModel", dict(nx=32, dt=7200.0, tavestart=0.0, taveint=7200.0, **QG_PHYS),
           5, True, True, 12),
    "qg_sparse": ("QGModel", dict(nx=32, dt=3600.0, tavestart=5 * 3600.0, taveint=3 * 3600.0,
                                  **QG_PHYS), 7, False, False, 20),
    "layered3": ("LayeredModel", dict(nx=32, tavestart=0.0, taveint=3600.0, **LAYERED3),
                 6, False, True, 12),
    "layered2": ("LayeredModel", dict(nx=32, nz=2, rd=15000.0, delta=0.25, H=[500., 2000.],
                                      U=[0.025, 0.0], V=[0.0, 0.0], rho=[1025.0, 1025.3],
                                      dt=3600.0, tavestart=0.0, taveint=3600.0),
                 8, True, False, 8),
    "bt": ("BTModel", dict(nx=32, beta=1.5e-11, rd=15000.0, H=1.0, U=0.01, dt=3600.0,
                           tavestart=0.0, taveint=3600.0), 9, False, False, 8),
    "sqg": ("SQGModel", dict(nx=32, beta=0.0, H=1.0, U=0.0, dt=3600.0, tavestart=0.0,
                             taveint=3600.0), 10, False, False, 8),
}


def q_param(m):
    return -2e-7 * m.q + 1e-13 * np.cos(2 * np.pi * m.x / m.L)[np.newaxis]


def uv_param_curl(m):
    return 1e-6 * m.v, 1e-6 * m.u + 1e-9 * np.cos(2 * np.pi * m.x / m.L)[np.newaxis]


def uv_param(m):
    return 1e-6 * m.v, -1e-6 * m.u + 1e-9 * np.sin(2 * np.pi * m.y / m.W)[np.newaxis]


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


def close(a, b, tol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = np.abs(b).max()
    if scale == 0.0:
        return np.abs(a).max() <= 1e-300
    return np.abs(a - b).max() <= tol * scale


def build(pq, case):
    cls, kw, seed, qp, uvp, steps = CASES[case]
    extra = {}
    if qp:
        extra["q_parameterization"] = q_param
    if uvp:
        extra["uv_parameterization"] = uv_param_curl
    m = getattr(pq, cls)(log_level=0, **kw, **extra)
    m.q = 1e-6 * np.random.default_rng(seed).standard_normal((m.nz, m.ny, m.nx))
    return m, steps


@pytest.mark.parametrize("case", list(CASES))
def test_diagnostics_match_reference(pq, case):
    print(f"diagnostics {case}", flush=True)
    g = golden("diagnostics.npz")
    m, steps = build(pq, case)
    for _ in range(steps):
        m._step_forward()
    expected = sorted(k[len(case) + 7:] for k in g.files if k.startswith(f"{case}_count_"))
    assert sorted(n for n, d in m.diagnostics.items() if d["count"]) == expected
    for n in expected:
        assert m.diagnostics[n]["count"] == int(g[f"{case}_count_{n}"]), n
        assert close(m.get_diagnostic(n), g[f"{case}_{n}"], 1e-9), n


def test_unfilled_diagnostic_raises(pq):
    from pyqg_amd.errors import DiagnosticNotFilledError
    m = pq.QGModel(nx=32, log_level=0)
    m.describe_diagnostics()
    with pytest.raises(DiagnosticNotFilledError):
        m.get_diagnostic("KEspec")
    assert "value" not in m.diagnostics["KEspec"]


def test_run_with_fused_stretches_matches_oracle(pq):
    """run(): fused steps between accumulation steps, phase-by-phase steps
    where diagnostics are due; same means as the oracle stepping every step."""
    from oracle.qg_oracle import OracleQG
    from oracle.qg_diagnostics import Accumulator, step_with_diagnostics
    dt = 3600.0
    kw = dict(nx=64, dt=dt, **QG_PHYS)
    q0 = 1e-6 * np.random.default_rng(11).standard_normal((2, 64, 64))
    m = pq.QGModel(log_level=0, tmax=40 * dt, tavestart=20 * dt, taveint=5 * dt, **kw)
    m.q = q0
    m.run()
    o = OracleQG(**kw)
    o.q = q0
    acc = Accumulator(o, tavestart=20 * dt, taveint=5 * dt)
    while o.t < 40 * dt:
        step_with_diagnostics(o, acc)
    # due at the starts of steps tc = 20, 25, 30, 35 (model.py:577-580)
    assert m.diagnostics["KEspec"]["count"] == acc.count == 4
    for n in acc.active:
        assert close(m.get_diagnostic(n), acc.get(n), 1e-9), n


def test_instant_functions_match_oracle(pq):
    """The registry's function entries evaluate the current fields
    (the form test_diagnostics.py:163-195 calls them in)."""
    from oracle.qg_oracle import OracleLayered
    from oracle.qg_diagnostics import attach_constants, derived_fields, evaluate
    kw = {k: v for k, v in LAYERED3.items()}
    q0 = 1e-6 * np.random.default_rng(12).standard_normal((3, 32, 32))
    m = pq.LayeredModel(nx=32, log_level=0, **kw)
    m.q = q0
    o = OracleLayered(nx=32, **kw)
    o.q = q0
    for _ in range(3):
        m._step_forward()
        o._step_forward()
    m._invert()
    m._calc_derived_fields()
    o._invert()
    attach_constants(o)
    derived_fields(o)
    for n, d in m.diagnostics.items():
        if n in ("Dissspec", "ENSDissspec", "paramspec", "ENSparamspec", "paramspec_APEflux",
                 "paramspec_KEflux"):
            continue  # need this step's tendency, which _invert alone does not compute
        assert close(d["function"](m), evaluate(o, n), 1e-9), n
    # host derived fields agree with the oracle's
    assert close(m.JSp, o.JSp, 1e-10) and close(m.phn, o.phn, 1e-10)


def test_paramspec_additivity(pq):
    """test_diagnostics.py:51-99: paramspec of the model with both
    parameterizations is the sum of the single ones; zero without any."""
    for cls in ("QGModel", "LayeredModel"):
        base = getattr(pq, cls)(nx=32, log_level=0, tavestart=0.0, taveint=1.0)
        rng = np.random.default_rng(3)
        dq = rng.normal(size=base.q.shape) * 1e-10
        du = rng.normal(size=base.q.shape) * 1e-6
        dv = rng.normal(size=base.q.shape) * 1e-6
        ms = [base,
              getattr(pq, cls)(nx=32, log_level=0, q_parameterization=lambda m: dq),
              getattr(pq, cls)(nx=32, log_level=0, uv_parameterization=lambda m: (du, dv)),
              getattr(pq, cls)(nx=32, log_level=0, q_parameterization=lambda m: dq,
                               uv_parameterization=lambda m: (du, dv))]
        q0 = 1e-6 * rng.standard_normal(base.q.shape)
        for m in ms:
            m.q = q0
            m._step_forward()
            m._increment_diagnostics()
        assert np.abs(ms[0].get_diagnostic("paramspec")).max() == 0.0
        for m in ms[1:]:
            assert np.abs(m.get_diagnostic("paramspec")).max() > 0.0
        np.testing.assert_allclose(ms[1].get_diagnostic("paramspec")
                                   + ms[2].get_diagnostic("paramspec"),
                                   ms[3].get_diagnostic("paramspec"), rtol=1e-10, atol=1e-30)


def test_dissspec_consistent_with_model_fields(pq):
    """test_diagnostics.py:101-154 on the device model."""
    dt = 3600.0
    m = pq.QGModel(nx=64, log_level=0, tavestart=100 * dt, taveint=dt, tmax=100 * dt, dt=dt)
    m.q = 1e-6 * np.random.default_rng(4).standard_normal((2, 64, 64))
    m.run()
    m._invert()
    m._do_advection()
    m._do_friction()
    m._calc_diagnostics()
    from pyqg_amd.constants import ab_weights
    (dt1, dt2, dt3), _ = ab_weights(m.ablevel, m.dt)
    rhs = m.qh + dt1 * m.dqhdt + dt2 * m.dqhdt_p + dt3 * m.dqhdt_pp
    diss = (m.filtr - 1.0) * rhs
    want = -np.real(np.tensordot(m.Hi, np.conj(m.ph) * diss, axes=(0, 0))) / m.H / m.dt / m.M ** 2
    assert m.diagnostics["Dissspec"]["count"] == 1
    np.testing.assert_allclose(m.get_diagnostic("Dissspec"), want, rtol=0,
                               atol=1e-9 * np.abs(want).max())


def test_ensemble_diagnostics_per_member(pq):
    """Ensemble.accumulate_diagnostics gives each member the means its own
    oracle model accumulates at the same states."""
    from oracle.qg_oracle import OracleQG
    from oracle.qg_diagnostics import Accumulator, names
    M, nx, dt = 3, 64, 3600.0
    betas = np.array([1.0e-11, 1.5e-11, 2.0e-11])
    ens = pq.Ensemble.qg(nx, beta=betas, dt=dt, nmembers=M, full_state=True)
    q0 = [1e-6 * np.random.default_rng(20 + i).standard_normal((2, nx, nx)) for i in range(M)]
    ens.set_q(q0)
    phys = {k: v for k, v in QG_PHYS.items() if k != "beta"}
    orc = []
    for i in range(M):
        o = OracleQG(nx=nx, dt=dt, beta=betas[i], **phys)
        o.q = q0[i]
        orc.append((o, Accumulator(o)))
    qg_names = names(orc[0][0])
    for rep in range(3):
        ens.step(7)
        ens.accumulate_diagnostics(qg_names)
        for o, acc in orc:
            for _ in range(7):
                o._step_forward()
            o._invert()
            o._do_advection()
            o._do_friction()
            acc.increment()
    for n in qg_names:
        for i, (o, acc) in enumerate(orc):
            assert close(ens.get_diagnostic(n, member=i), acc.get(n), 1e-9), (n, i)
    mean = ens.get_diagnostic("KEspec", mean=True)[0]
    assert close(mean, np.mean([acc.get("KEspec") for _, acc in orc], axis=0), 1e-9)


@pytest.mark.parametrize("nz", [16, 17])
def test_many_layers_run_and_diagnostics_match_oracle(pq, nz):
    """nz = 16 is the last layer count of the column kernel that keeps ph in
    its own buffer, nz = 17 runs the kernel for any nz (and the diagnostics'
    larger per-point arrays): fused stretches plus accumulation steps against
    the oracle stepping every step."""
    from oracle.qg_oracle import OracleLayered
    from oracle.qg_diagnostics import Accumulator, step_with_diagnostics
    from pyqg_amd.constants import exponential_stratification
    print(f"layered nz={nz} run + diagnostics", flush=True)
    dt = 3600.0
    H, rho, U = exponential_stratification(nz=nz)
    kw = dict(nz=nz, nx=32, H=H, rho=rho, U=U, V=np.zeros(nz), L=1e6, beta=1.5e-11,
              rek=5.787e-7, dt=dt, f=0.0001236812857687059, g=9.81)
    q0 = 1e-7 * np.random.default_rng(nz).standard_normal((nz, 32, 32))
    m = pq.LayeredModel(log_level=0, tmax=24 * dt, tavestart=12 * dt, taveint=4 * dt, **kw)
    m.q = q0
    m.run()
    o = OracleLayered(**kw)
    o.q = q0
    acc = Accumulator(o, tavestart=12 * dt, taveint=4 * dt)
    while o.t < 24 * dt:
        step_with_diagnostics(o, acc)
    assert rel_l2(m.q, o.q) < 1e-10
    assert m.diagnostics["KEspec"]["count"] == acc.count > 0
    for n in acc.active:
        assert close(m.get_diagnostic(n), acc.get(n), 1e-9), n
