"""The measured ensemble workloads against the CPU oracle, with the kernels
they actually dispatch.

The headline workload (256 two-layer members at nx=256) and the uneven and
layered ensembles run the large-bucket kernels: the radix-8 register row
kernel and the two-layer register column kernel, or the nz >= 3 column kernel
with the tridiagonal solve.  These tests step exactly those buckets (default
variant, every radix, and the autotuner's choice) and compare members against
the oracle at the 1e-10 trajectory bar.
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


@pytest.fixture(scope="module")
def headline_oracle():
    """q after 100 steps of members 0, 37 and 255 of the headline ensemble."""
    out = {}
    for i in (0, 37, 255):
        o = OracleQG(nx=256, dt=3600., **QG_PHYS)
        o.q = seeded_q(2, 256, i)
        for _ in range(100):
            o._step_forward()
        out[i] = o.q.copy()
    return out


def headline_ensemble(pq):
    ens = pq.Ensemble.qg(np.full(256, 256), dt=3600.)
    ens.set_q([seeded_q(2, 256, i) for i in range(256)])
    return ens


@pytest.mark.parametrize("variant", ["default", (4, 4), (16, 16), (4, 16), (16, 8)])
def test_headline_ensemble_matches_oracle(pq, headline_oracle, variant):
    """256 members of nx=256 (the bench's bucket): 100 fused steps with the
    register column kernel at every radix (the default is radix 8 for both
    kernels), members 0, 37 and 255 against the oracle."""
    print(f"headline ensemble variant {variant}", flush=True)
    ens = headline_ensemble(pq)
    b = ens.buckets[0]
    assert not b.layer_split()
    if variant == "default":
        assert b.variant() == (8, 8)
    else:
        b.set_variant(*variant)
    ens.step(100)
    for i, ref in headline_oracle.items():
        assert rel_l2(ens.get("q", i), ref) < TOL_TRAJ, (variant, i)


def test_headline_ensemble_autotuned_graph(pq, headline_oracle):
    """As the bench runs it: three start-up steps, the autotuner (which times
    every radix on the live state and restores it), then graph replays."""
    print("headline ensemble autotuned + graph", flush=True)
    ens = headline_ensemble(pq)
    ens.step(3)
    report = ens.autotune(reps=2, cache_path=False)
    assert len(report) == 1
    ens.step(97, use_graph=True)
    for i, ref in headline_oracle.items():
        assert rel_l2(ens.get("q", i), ref) < TOL_TRAJ, i


def test_uneven_ensemble_as_benchmarked(pq):
    """The uneven workload exactly as the bench builds it (256 members, nx from
    {64, 128, 256}, per-member beta, rd, delta, rek, default_rng(1)),
    autotuned; one member of each grid size against the oracle."""
    print("uneven 256-member ensemble", flush=True)
    M = 256
    rng = np.random.default_rng(1)
    nxs = rng.choice([64, 128, 256], size=M)
    beta = rng.uniform(0.5e-11, 2.5e-11, M)
    rd = rng.uniform(1e4, 2.5e4, M)
    delta = rng.uniform(0.1, 0.5, M)
    rek = rng.uniform(3e-7, 1e-6, M)
    ens = pq.Ensemble.qg(nxs, beta=beta, rd=rd, delta=delta, rek=rek, dt=3600.)
    assert sorted(b.nx for b in ens.buckets) == [64, 128, 256]
    q0 = [seeded_q(2, int(n), s) for s, n in enumerate(nxs)]
    ens.set_q(q0)
    nsteps = 30
    ens.step(3)
    ens.autotune(reps=2, cache_path=False)
    ens.step(nsteps - 3)
    picks = [int(np.nonzero(nxs == n)[0][-1]) for n in (64, 128, 256)]
    for i in picks:
        o = OracleQG(nx=int(nxs[i]), dt=3600., beta=beta[i], rd=rd[i], delta=delta[i],
                     rek=rek[i], L=1e6, H1=500., U1=0.025, U2=0.0, filterfac=23.6)
        o.q = q0[i]
        for _ in range(nsteps):
            o._step_forward()
        assert rel_l2(ens.get("q", i), o.q) < TOL_TRAJ, (i, int(nxs[i]))


@pytest.mark.parametrize("variant", ["default", (4, 4), (16, 16)])
def test_layered_nz10_nx256_ensemble(pq, variant):
    """64 LayeredModel members, nz=10, nx=256 (the layered workload): the
    nz >= 3 column kernel with the tridiagonal solve, 20 steps, two members
    against the oracle (layered_model.py:211-229 inversion)."""
    print(f"layered nz=10 nx=256 ensemble variant {variant}", flush=True)
    H, rho, U = pq.constants.exponential_stratification()
    nx, n, nsteps = 256, 64, 20
    ens = pq.Ensemble.layered(n, nx, 10, H, rho, U, L=2e6, dt=3600.)
    b = ens.buckets[0]
    assert b.inversion_tridiagonal()
    if variant == "default":
        assert b.variant() == (8, 8)
    else:
        b.set_variant(*variant)
    q0 = [seeded_q(10, nx, 500 + i) for i in range(n)]
    ens.set_q(q0)
    ens.step(nsteps)
    kw = dict(nz=10, nx=nx, H=H, rho=rho, U=U, V=np.zeros(10), L=2e6, beta=1.5e-11,
              rek=5.787e-7, dt=3600.0, f=0.0001236812857687059, g=9.81)
    for i in (0, n - 1):
        o = OracleLayered(**kw)
        o.q = q0[i]
        for _ in range(nsteps):
            o._step_forward()
        assert rel_l2(ens.get("q", i), o.q) < TOL_TRAJ, (variant, i)


def test_member_health_check(pq):
    """The per-member CFL check agrees with the reference's formula
    (|u + Ubg|, |v| max times dt/dx) evaluated on the oracle's state, flags a
    member whose state is not finite, and masks it without touching the
    others."""
    ens = pq.Ensemble.qg(np.array([64, 64, 128, 64]), dt=3600.)
    q0 = [seeded_q(2, int(n), 70 + i) for i, n in enumerate([64, 64, 128, 64])]
    q0[1] = q0[1] * 1e3  # a fast member
    ens.set_q(q0)
    ens.step(12)
    cfl = ens.member_cfl()
    for i, n in enumerate([64, 64, 128, 64]):
        o = OracleQG(nx=n, dt=3600., **QG_PHYS)
        o.q = q0[i]
        for _ in range(12):
            o._step_forward()
        o._invert()           # u, v of the current state
        ref = np.abs(np.hstack([o.u + o.Ubg[:, None, None], o.v])).max() * 3600. / (1e6 / n)
        assert abs(cfl[i] - ref) <= 1e-9 * ref, (i, cfl[i], ref)
    assert ens.check_members() == []
    # a non-finite member: NaN into one point of member 3's PV
    bi, j = ens.members[3]
    b = ens.buckets[bi]
    qh = b.get("qh")
    qh[j, 1, 3, 5] = np.nan
    b.put("qh", qh)
    cfl = ens.member_cfl()
    assert not np.isfinite(cfl[3]) and np.isfinite(cfl[[0, 1, 2]]).all()
    with pytest.raises(AssertionError):
        ens.check_members(on_failure="raise")
    assert ens.check_members() == [3] and ens.failed == {3}
    before = [ens.get("qh", i) for i in (0, 1, 2)]
    ens.step(4, twrite=2)
    assert not np.any(ens.get("qh", 3))
    # a masked member rests at q = 0: its only speed is the upper layer's U1
    assert np.isfinite(ens.cfl).all()
    assert abs(ens.cfl[3] - 0.025 * 3600. / (1e6 / 64)) < 1e-12
    ref = pq.Ensemble.qg(np.array([64, 64, 128, 64]), dt=3600.)
    ref.set_q(q0)
    ref.step(16)
    for i in (0, 1, 2):
        assert rel_l2(ens.get("qh", i), ref.get("qh", i)) < 1e-13, i
    assert not np.array_equal(before[0], ens.get("qh", 0))
