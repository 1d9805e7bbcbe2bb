"""GPU trajectories against the CPU oracle on identical seeded inputs.

The bar: relative L2 of q <= 1e-10 after 100 steps (before chaotic
divergence), per-phase fields <= 1e-12, at sizes the oracle finishes in
seconds; the nx=512, dt=1800 two-layer run is the single-GPU parity workload.
"""
import numpy as np
import pytest

from conftest import QG_PHYS, rel_l2
from oracle.qg_oracle import OracleLayered, OracleQG

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-10


@pytest.fixture(scope="module")
def pq(gpu):
    import pyqg_amd
    return pyqg_amd


def seeded_q(nz, n, seed):
    return 1e-7 * np.random.default_rng(seed).standard_normal((nz, n, n))


def test_phases_match_oracle_on_turbulent_state(pq):
    """Each reference method on the spun-up LayeredModel fixture state."""
    from conftest import golden
    g = golden("layered_nz3_nx64.npz")
    kw = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
              rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
              beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)
    m = pq.LayeredModel(nx=64, log_level=0, **kw)
    o = OracleLayered(nx=64, **kw)
    o.a = np.ascontiguousarray(m.a.real)   # same constants on both sides
    m.q = g["q0"]
    o.q = g["q0"]
    for step in range(3):
        for phase, fields in (("_invert", ["ph", "uh", "vh", "u", "v"]),
                              ("_do_advection", ["uq", "vq", "uqh", "vqh", "dqhdt"]),
                              ("_do_friction", ["dqhdt"]),
                              ("_forward_timestep", ["qh", "q", "dqhdt_p", "dqhdt_pp"])):
            getattr(m, phase)()
            getattr(o, phase)()
            for f in fields:
                ref = getattr(o, f)
                if np.abs(ref).max() == 0:
                    continue
                assert rel_l2(getattr(m, f), ref) < 1e-12, (step, phase, f)


@pytest.mark.parametrize("path", ["methods", "fused"])
def test_qg_nx64_100_steps(pq, path):
    m = pq.QGModel(nx=64, dt=7200.0, log_level=0, **QG_PHYS)
    o = OracleQG(nx=64, dt=7200.0, **QG_PHYS)
    q0 = seeded_q(2, 64, 0)
    m.q = q0
    o.q = q0
    for _ in range(100):
        o._step_forward()
    if path == "methods":
        for _ in range(100):
            m._step_forward()
    else:
        m._fused_steps(100)
    assert m.t == o.t and m.tc == 100
    assert rel_l2(m.q, o.q) < TOL_TRAJ
    assert rel_l2(m.dqhdt_p, o.dqhdt_p) < TOL_TRAJ
    assert rel_l2(m.ph, o.ph) < TOL_TRAJ


def test_qg_nx512_dt1800_config(pq):
    """Two-layer nx=512, dt=1800, seeded 1e-7 Gaussian q: 1e-10 at step 100."""
    print("nx=512 fused 100 steps", flush=True)
    m = pq.QGModel(nx=512, dt=1800.0, log_level=0, **QG_PHYS)
    o = OracleQG(nx=512, dt=1800.0, **QG_PHYS)
    q0 = seeded_q(2, 512, 0)
    m.q = q0
    o.q = q0
    m._fused_steps(100)
    for _ in range(100):
        o._step_forward()
    err = rel_l2(m.q, o.q)
    print(f"nx=512 step-100 rel L2 = {err:.3e}", flush=True)
    assert err < TOL_TRAJ


def test_uneven_qg_ensemble_matches_oracle(pq):
    """Members of nx in {64,128,256} with their own beta, rd, delta, rek."""
    rng = np.random.default_rng(1)
    n = 6
    nx = np.array([64, 128, 256, 64, 128, 256])
    beta = rng.uniform(0.5e-11, 2.5e-11, n)
    rd = rng.uniform(1e4, 2.5e4, n)
    delta = rng.uniform(0.1, 0.5, n)
    rek = rng.uniform(3e-7, 1e-6, n)
    ens = pq.Ensemble.qg(nx, beta=beta, rd=rd, delta=delta, rek=rek, dt=3600.)
    q0 = [seeded_q(2, int(nx[i]), 100 + i) for i in range(n)]
    ens.set_q(q0)
    nsteps = 30
    ens.step(nsteps)
    for i in range(n):
        o = OracleQG(nx=int(nx[i]), dt=3600., beta=beta[i], rd=rd[i], delta=delta[i],
                     rek=rek[i], L=1e6, H1=500., U1=0.025, U2=0.0, filterfac=23.6)
        o.q = q0[i]
        for _ in range(nsteps):
            o._step_forward()
        assert rel_l2(ens.get("q", i), o.q) < TOL_TRAJ, i
        assert rel_l2(ens.get("qh", i), o.qh) < TOL_TRAJ, i
    assert ens.tc == nsteps


def test_ensemble_chunked_steps_equal_single_call(pq):
    ens1 = pq.Ensemble.qg(np.full(4, 64), dt=3600.)
    ens2 = pq.Ensemble.qg(np.full(4, 64), dt=3600.)
    q0 = [seeded_q(2, 64, 7 + i) for i in range(4)]
    ens1.set_q(q0)
    ens2.set_q(q0)
    ens1.step(10)
    for _ in range(5):
        ens2.step(2)
    for i in range(4):
        assert np.array_equal(ens1.get("qh", i), ens2.get("qh", i))


def test_layered_nz10_ensemble_shared_inversion(pq):
    H, rho, U = pq.constants.exponential_stratification()
    nx, n, nsteps = 64, 3, 20
    ens = pq.Ensemble.layered(n, nx, 10, H, rho, U, L=2e6, dt=3600.)
    q0 = [seeded_q(10, nx, 30 + i) for i in range(n)]
    ens.set_q(q0)
    ens.step(nsteps)
    o = OracleLayered(nz=10, nx=nx, H=H, rho=rho, U=U, V=np.zeros(10), L=2e6, beta=1.5e-11,
                      rek=5.787e-7, dt=3600.0, f=0.0001236812857687059, g=9.81)
    np.testing.assert_allclose(ens.a, o.a.real, rtol=1e-9, atol=1e-9 * np.abs(o.a).max())
    for i in range(n):
        o2 = OracleLayered(nz=10, nx=nx, H=H, rho=rho, U=U, V=np.zeros(10), L=2e6,
                           beta=1.5e-11, rek=5.787e-7, dt=3600.0, f=0.0001236812857687059,
                           g=9.81)
        o2.a = ens.a
        o2.q = q0[i]
        for _ in range(nsteps):
            o2._step_forward()
        assert rel_l2(ens.get("q", i), o2.q) < TOL_TRAJ, i


def test_layered_tridiagonal_solve_matches_table(pq):
    """Fused nz=10 steps solving (S - kappa^2 I) ph = qh agree with the same
    steps reading the inversion table a, and with the oracle; writing a
    switches the solve off."""
    H, rho, U = pq.constants.exponential_stratification()
    kw = dict(nz=10, nx=64, H=H, rho=rho, U=U, V=np.zeros(10), L=2e6, beta=1.5e-11,
              rek=5.787e-7, dt=3600.0, f=0.0001236812857687059, g=9.81)
    m1 = pq.LayeredModel(log_level=0, **kw)
    m2 = pq.LayeredModel(log_level=0, **kw)
    assert m1._bucket.inversion_tridiagonal() and m2._bucket.inversion_tridiagonal()
    m2._bucket.set_inversion_tridiagonal(None)
    assert not m2._bucket.inversion_tridiagonal()
    o = OracleLayered(**kw)
    o.a = np.ascontiguousarray(m1.a.real)
    q0 = seeded_q(10, 64, 77)
    for m in (m1, m2, o):
        m.q = q0
    m1._fused_steps(60)
    m2._fused_steps(60)
    for _ in range(60):
        o._step_forward()
    assert rel_l2(m1.q, m2.q) < 1e-12
    assert rel_l2(m1.ph, m2.ph) < 1e-12
    assert rel_l2(m1.q, o.q) < TOL_TRAJ
    m1.a = m1.a.real.copy()          # a user-written table is used as given
    assert not m1._bucket.inversion_tridiagonal()
    with pytest.raises(Exception):   # two layers: no tridiagonal solve
        pq.QGModel(nx=32, log_level=0)._bucket.set_inversion_tridiagonal(np.eye(2))


def test_full_size_invariants_headline_ensemble(pq):
    """256 members of nx=256 (the measured workload): members with identical
    inputs stay bit-identical, and the state stays finite and bounded."""
    ens = pq.Ensemble.qg(np.full(256, 256), dt=3600.)
    q0 = [seeded_q(2, 256, i % 8) for i in range(256)]
    ens.set_q(q0)
    ens.step(20)
    qh = ens.get("qh")
    for i in range(8, 256, 37):
        assert np.array_equal(qh[i], qh[i % 8])
    assert all(np.isfinite(x).all() for x in qh)


@pytest.mark.parametrize("erow,ecol", [(4, 4), (4, 8), (8, 16), (16, 4), (16, 16), (8, 8)])
def test_fused_radix_variants(pq, erow, ecol):
    """Every register radix of the fused kernels gives the oracle's trajectory:
    the two-layer register column kernel (layer split off) and the layer-split
    one, the one-layer register path and the generic nz path."""
    nsteps = 12
    q0 = [seeded_q(2, n, 40 + i) for i, n in enumerate([32, 64, 64])]
    refs = []
    for i, n in enumerate([32, 64, 64]):
        o = OracleQG(nx=n, dt=3600., **QG_PHYS)
        o.q = q0[i]
        for _ in range(nsteps):
            o._step_forward()
        refs.append(o.q.copy())
    for split in (True, False):
        ens = pq.Ensemble.qg(np.array([32, 64, 64]), dt=3600.)
        for b in ens.buckets:
            b.set_variant(erow, ecol)
            b.set_layer_split(split)
            assert b.layer_split() == split
        ens.set_q(q0)
        ens.step(nsteps)
        for i in range(3):
            assert rel_l2(ens.get("q", i), refs[i]) < TOL_TRAJ, (split, i)
    # one two-layer member: both column kernels at this radix
    for split in (True, False):
        qm = pq.QGModel(nx=64, dt=3600., log_level=0, **QG_PHYS)
        qm._bucket.set_variant(erow, ecol)
        qm._bucket.set_layer_split(split)
        qm.q = q0[1]
        qm._fused_steps(nsteps)
        assert rel_l2(qm.q, refs[1]) < TOL_TRAJ, split
    from oracle.qg_oracle import OracleBT
    bt = pq.BTModel(L=2 * np.pi, nx=64, beta=0.5, rd=0.3, dt=0.005, log_level=0)
    bt._bucket.set_variant(erow, ecol)
    ob = OracleBT(L=2 * np.pi, nx=64, beta=0.5, rd=0.3, dt=0.005)
    qb = 1e-2 * np.random.default_rng(3).standard_normal((1, 64, 64))
    bt.q = qb
    ob.q = qb
    bt._fused_steps(nsteps)
    for _ in range(nsteps):
        ob._step_forward()
    assert rel_l2(bt.q, ob.q) < TOL_TRAJ
    kw = dict(nz=3, U=[0.05, 0.025, 0.0], V=[0.0, 0.0, 0.0], H=[500, 1750, 1750],
              rho=[1025.0, 1025.3, 1025.6], L=1e6, rek=5.787e-7, g=9.81, dt=3600.0,
              beta=1.5e-11, rd=15000.0, f=0.0001236812857687059)
    lm = pq.LayeredModel(nx=32, log_level=0, **kw)
    lm._bucket.set_variant(erow, ecol)
    ol = OracleLayered(nx=32, **kw)
    ol.a = np.ascontiguousarray(lm.a.real)
    ql = seeded_q(3, 32, 9)
    lm.q = ql
    ol.q = ql
    lm._fused_steps(nsteps)
    for _ in range(nsteps):
        ol._step_forward()
    assert rel_l2(lm.q, ol.q) < TOL_TRAJ


@pytest.mark.parametrize("n,erow,ecol,split", [(64, 8, 8, False), (64, 4, 4, True),
                                                (256, 8, 8, False), (48, 8, 8, False)])
def test_fused_nyquist_column_unfiltered(pq, n, erow, ecol, split):
    """With the exponential filter off (filterfac = 0) the Nyquist column
    kx = nx/2 keeps its full amplitude (the default filter scales it by ~1e-15
    every step, so q alone cannot show it).  After fused steps qh and both
    tendency histories match the oracle over the whole field, and in column 0
    and the Nyquist column on their own: the register column kernels carry
    these two columns in one packed slot of the x-mixed exchange."""
    phys = dict(QG_PHYS, filterfac=0.0)
    nsteps = 8
    q0 = seeded_q(2, n, 71)
    m = pq.QGModel(nx=n, dt=3600., log_level=0, **phys)
    m._bucket.set_variant(erow, ecol)
    m._bucket.set_layer_split(split)
    o = OracleQG(nx=n, dt=3600., **phys)
    m.q = q0
    o.q = q0
    m._fused_steps(nsteps)
    for _ in range(nsteps):
        o._step_forward()
    for name in ("qh", "dqhdt_p", "dqhdt_pp"):
        a, b = np.asarray(getattr(m, name)), np.asarray(getattr(o, name))
        assert np.abs(b[..., -1]).max() > 1e-6 * np.abs(b).max(), name  # Nyquist not negligible
        assert rel_l2(a, b) < 1e-11, name
        assert rel_l2(a[..., 0], b[..., 0]) < 1e-11, name
        assert rel_l2(a[..., -1], b[..., -1]) < 1e-11, name
    assert rel_l2(m.q, o.q) < 1e-11


def test_packed_exchange_switch(pq):
    """The register column kernels exchange nx/2 x-mixed columns per row
    (bins 0 and nx/2 packed into one slot); switched off they exchange
    nx/2 + 1.  Both give the same trajectory to rounding, and the oracle's,
    with the filter off so that the Nyquist column carries weight."""
    phys = dict(QG_PHYS, filterfac=0.0)
    q0 = [seeded_q(2, 128, 80 + i) for i in range(40)]
    runs = []
    for on in (True, False):
        ens = pq.Ensemble.qg(np.full(40, 128), dt=3600., **phys)
        b = ens.buckets[0]
        assert not b.layer_split()
        b.set_packed_exchange(on)
        assert b.exchange_columns() == (64 if on else 65)
        ens.set_q(q0)
        ens.step(10)
        runs.append([np.asarray(ens.get("qh", i)) for i in (0, 39)])
    o = OracleQG(nx=128, dt=3600., **phys)
    o.q = q0[39]
    for _ in range(10):
        o._step_forward()
    for a, b in zip(*runs):
        assert rel_l2(a, b) < 1e-12
        assert rel_l2(a[..., -1], b[..., -1]) < 1e-12
    assert rel_l2(runs[0][1], np.asarray(o.qh)) < 1e-11
    assert rel_l2(runs[0][1][..., -1], np.asarray(o.qh)[..., -1]) < 1e-11
    split = pq.Ensemble.qg(np.full(2, 64), dt=3600.).buckets[0]
    assert split.layer_split() and split.exchange_columns() == 33


def test_nonstandard_wavenumber_tables_use_the_table_kernels(pq):
    """The packed kernels form ll and kappa^2 themselves when every member's
    tables are the standard ones.  A kappa^2 table that differs from
    kk^2 + ll^2 anywhere (here by one ulp at one point, through the drop-in
    _k2l2 setter) keeps the table-reading kernels, whose trajectory then
    agrees with the standard run's to rounding."""
    n, nsteps = 64, 12
    q0 = seeded_q(2, n, 95)
    runs = []
    for perturb in (False, True):
        m = pq.QGModel(nx=n, dt=3600., log_level=0, **QG_PHYS)
        m._bucket.set_variant(8, 8)
        m._bucket.set_layer_split(False)
        assert m._bucket.exchange_columns() == n // 2
        if perturb:
            k2 = np.array(m._k2l2)
            k2[5, 7] = np.nextafter(k2[5, 7], np.inf)
            m._k2l2 = k2
            assert m._bucket.exchange_columns() == n // 2 + 1
        m.q = q0
        m._fused_steps(nsteps)
        runs.append(np.array(m.q))
    assert rel_l2(runs[1], runs[0]) < 1e-12


@pytest.mark.parametrize("n", [128, 256, 1024, 384, 768])
def test_single_member_default_radix(pq, n):
    """A one-member model steps its fused stretches with radix 4 (radix 6 for
    3 * 2^a): the large power-of-two and mixed-radix grids against the oracle."""
    print(f"single-member fused n={n}", flush=True)
    nsteps = 10
    m = pq.QGModel(nx=n, dt=3600., tmax=nsteps * 3600., log_level=0, **QG_PHYS)
    assert m._bucket.variant() == (4, 4)
    o = OracleQG(nx=n, dt=3600., **QG_PHYS)
    q0 = seeded_q(2, n, 60)
    m.q = q0
    o.q = q0
    m.run()
    for _ in range(nsteps):
        o._step_forward()
    assert m.tc == nsteps
    assert rel_l2(m.q, o.q) < 1e-12


def test_small_bucket_defaults_and_layer_split(pq):
    """Buckets with members x n^2 <= 2^19 (and single members) default to
    radix 4 and the layer-split column kernel, larger ones to radix 8 and the
    register kernel; both column kernels give the same trajectory to rounding,
    and the oracle's."""
    small = pq.Ensemble.qg(np.full(4, 64), dt=3600.)
    large = pq.Ensemble.qg(np.full(40, 128), dt=3600.)
    assert small.buckets[0].variant() == (4, 4) and small.buckets[0].layer_split()
    assert large.buckets[0].variant() == (8, 8) and not large.buckets[0].layer_split()
    other = pq.Ensemble.qg(np.full(4, 64), dt=3600.)
    other.buckets[0].set_layer_split(False)
    q0 = [seeded_q(2, 64, 90 + i) for i in range(4)]
    small.set_q(q0)
    other.set_q(q0)
    small.step(40)
    other.step(40)
    o = OracleQG(nx=64, dt=3600., **QG_PHYS)
    o.q = q0[2]
    for _ in range(40):
        o._step_forward()
    for i in range(4):
        assert rel_l2(small.get("q", i), other.get("q", i)) < 1e-12
    assert rel_l2(small.get("q", 2), o.q) < TOL_TRAJ
    with pytest.raises(Exception):
        pq.BTModel(nx=32, log_level=0)._bucket.set_layer_split(True)


def test_autotune_keeps_state_and_parity(pq, tmp_path):
    ens = pq.Ensemble.qg(np.array([64, 64, 128]), dt=3600.)
    q0 = [seeded_q(2, n, 60 + i) for i, n in enumerate([64, 64, 128])]
    ens.set_q(q0)
    ens.step(3)
    before = [ens.get(f) for f in ("qh", "dqhdt_p", "dqhdt_pp")]
    report = ens.autotune(reps=2, cache_path=str(tmp_path / "tune.json"))
    assert len(report) == 2
    after = [ens.get(f) for f in ("qh", "dqhdt_p", "dqhdt_pp")]
    for x, y in zip(before, after):
        for a, b in zip(x, y):
            assert np.array_equal(a, b)
    ens.step(7)
    for i, n in enumerate([64, 64, 128]):
        o = OracleQG(nx=n, dt=3600., **QG_PHYS)
        o.q = q0[i]
        for _ in range(10):
            o._step_forward()
        assert rel_l2(ens.get("q", i), o.q) < TOL_TRAJ
    # a second ensemble of the same shapes takes the variant from the cache
    ens2 = pq.Ensemble.qg(np.array([64, 64, 128]), dt=3600.)
    rep2 = ens2.autotune(cache_path=str(tmp_path / "tune.json"))
    assert [b.variant() for b in ens2.buckets] == [b.variant() for b in ens.buckets]
    assert rep2.keys() == report.keys()


def test_graph_replays_match_eager_steps(pq):
    """hipGraph replay of the fused step (two steps per replay, buckets of
    several sizes on one stream) is bit-identical to eager stepping, replayed
    many times with eager work and host synchronisation between replays."""
    nx = np.array([64, 128, 64, 32])
    q0 = [seeded_q(2, int(n), 80 + i) for i, n in enumerate(nx)]
    eager = pq.Ensemble.qg(nx, dt=3600.)
    graph = pq.Ensemble.qg(nx, dt=3600.)
    eager.set_q(q0)
    graph.set_q(q0)
    for chunk in (3, 6, 5, 1, 8, 2, 7):
        eager.step(chunk)
        graph.step(chunk, use_graph=True)
        for i in range(len(nx)):
            assert np.array_equal(eager.get("qh", i), graph.get("qh", i)), (chunk, i)
            assert np.array_equal(eager.get("dqhdt_p", i), graph.get("dqhdt_p", i))
    assert graph.tc == eager.tc == 32
    # changing a constant between replays takes effect (parameters live on the device)
    for e in (eager, graph):
        for b in e.buckets:
            b.set_rek(2e-6)
    eager.step(4)
    graph.step(4, use_graph=True)
    for i in range(len(nx)):
        assert np.array_equal(eager.get("qh", i), graph.get("qh", i))
