"""Parameterized runs replayed from HIP graphs.

Model.run() with built-in closures replays a captured graph of two
phase-by-phase steps between the diagnostic and status steps.  The graph holds
the same kernels with the same arguments as _step_forward, so the state and
the averaged diagnostics match a plain _step_forward loop bit for bit.  Every
test replays the graph several times with ordinary steps, host reads and
synchronizations in between."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DT = 7200.0


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


CLOSURES = {
    "smag": lambda pq: pq.Smagorinsky(0.15),
    "back": lambda pq: pq.BackscatterBiharmonic(0.08, 1.1),
    "zb": lambda pq: pq.ZannaBolton2020(),
    "combo": lambda pq: 0.5 * pq.Smagorinsky() + 0.75 * pq.ZannaBolton2020(),
}


def make(pq, which, nx=32, nsteps=40):
    p = CLOSURES[which](pq)
    m = pq.QGModel(nx=nx, dt=DT, tmax=nsteps * DT, tavestart=10 * DT, taveint=7 * DT,
                   log_level=0, parameterization=p)
    m.q = 1e-6 * np.random.default_rng(11).standard_normal((2, nx, nx))
    return m


def stepped(m):
    while m.t < m.tmax:
        m._step_forward()


@pytest.mark.parametrize("which", list(CLOSURES))
def test_graph_run_matches_step_forward(pq, which):
    print(f"graph run {which}", flush=True)
    a, b = make(pq, which), make(pq, which)
    assert a._graphable()
    for chunk in range(3):      # three run() calls: the cached graph is replayed again
        a.run()
        stepped(b)
        assert a.tc == b.tc and a.t == b.t
        np.testing.assert_array_equal(a.q, b.q)
        np.testing.assert_array_equal(a.qh, b.qh)
        for name in ("KEspec", "EKE", "Ensspec"):
            np.testing.assert_array_equal(a.get_diagnostic(name), b.get_diagnostic(name))
        a.tmax += 17 * DT       # 17 steps: the next stretches start at the other phase
        b.tmax += 17 * DT


def test_graph_follows_constant_changes(pq):
    """A new closure constant or a new drag coefficient between runs: the
    graph is recaptured (closure) or refused as stale and rebuilt (constants)."""
    a, b = make(pq, "smag"), make(pq, "smag")
    a.run()
    stepped(b)
    for m in (a, b):
        m.uv_parameterization.constant = 0.2
        m.tmax += 12 * DT
    a.run()
    stepped(b)
    np.testing.assert_array_equal(a.q, b.q)
    for m in (a, b):
        m.rek = 2 * m.rek
        m.tmax += 12 * DT
    a.run()
    stepped(b)
    np.testing.assert_array_equal(a.q, b.q)


def test_host_callback_closure_is_not_graphed(pq):
    """A parameterization evaluated on the host cannot be replayed: run()
    steps it one _step_forward at a time."""
    smag = pq.Smagorinsky(0.12)
    m = pq.QGModel(nx=32, dt=DT, tmax=6 * DT, log_level=0,
                   uv_parameterization=lambda mm: smag(mm))
    assert not m._graphable()
    m.q = 1e-6 * np.random.default_rng(3).standard_normal((2, 32, 32))
    m.run()
    assert m.tc == 6 and np.isfinite(m.q).all()


def test_capture_rejects_partial_history_cycle(pq):
    """A graph of an odd number of steps is refused (the tendency history's
    buffers would be left swapped)."""
    from pyqg_amd import _lib
    m = make(pq, "smag")
    for _ in range(3):
        m._step_forward()
    b = m._bucket
    L = _lib.lib
    assert L.qgk_capture_begin(b.h) == 0
    for fn in (L.qgk_invert, L.qgk_advect, L.qgk_friction):
        assert fn(b.h) == 0
    assert L.qgk_timestep(b.h, 1.0, 0.0, 0.0) == 0
    assert L.qgk_capture_end(b.h, 1) != 0
    assert "multiple of 2" in _lib.lib.qgk_last_error().decode()
    assert L.qgk_graph_replay(b.h, 1) != 0


@pytest.mark.parametrize("which", ["smag", "back"])
def test_ensemble_graph_steps_match_phase_by_phase(pq, which):
    """Ensemble.step_parameterized replays per-bucket graphs; in uneven chunks
    (odd lengths re-align the history phase) it matches an ensemble stepped
    phase by phase bit for bit."""
    from conftest import QG_PHYS
    print(f"ensemble graph {which}", flush=True)
    p = CLOSURES[which](pq)
    uv = p.parameterization_type == "uv_parameterization"
    phys = {k: v for k, v in QG_PHYS.items() if k != "beta"}
    betas = np.array([1.0e-11, 1.5e-11, 2.0e-11])
    ens = [pq.Ensemble.qg(32, beta=betas, dt=DT, nmembers=3, full_state=True,
                          q_param=not uv, uv_param=uv, **phys) for _ in range(2)]
    q0 = [1e-6 * np.random.default_rng(50 + i).standard_normal((2, 32, 32)) for i in range(3)]
    for e in ens:
        e.set_q(q0)
    for chunk in (7, 5, 6):
        ens[0].step_parameterized(chunk, p)
        for _ in range(chunk):
            ens[1]._param_step(p)
        assert ens[0].tc == ens[1].tc and ens[0].t == ens[1].t
        for i in range(3):
            np.testing.assert_array_equal(ens[0].get("q", i), ens[1].get("q", i))


def test_graph_recaptured_after_fused_steps(pq):
    """Fused steps swap the qh buffers: a graph captured before them is refused
    and rebuilt, so a model that drops its closure for a while and takes it
    back still steps exactly as _step_forward does."""
    a, b = make(pq, "smag", nsteps=12), make(pq, "smag", nsteps=12)
    a.run()
    stepped(b)
    smag = a.uv_parameterization
    for m in (a, b):
        m.uv_parameterization = None
        m.tmax += 8 * DT
        m.run()                 # fused stretches (an odd number of qh swaps) in both
        m.uv_parameterization = smag
        m.tmax += 10 * DT
    a.run()
    stepped(b)
    np.testing.assert_array_equal(a.q, b.q)
