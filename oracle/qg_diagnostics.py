"""CPU oracle of pyqg's time-averaged diagnostics.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker of the device
diagnostics; the package never imports it.

Restates, for the oracle models of qg_oracle.py:
  * the core registry of pyqg/model.py:644-736 (Ensspec, KEspec, EKEdiss,
    KEfrictionspec, EKE, Dissspec, ENSDissspec, paramspec, ENSparamspec) with
    _calc_parameterization_contribution / _spectrum (model.py:738-757);
  * QGDiagnostics (qg_diagnostics.py:6-70): entspec, paramspec_APEflux,
    paramspec_KEflux, ENSflux, ENSgenspec, ENSfrictionspec, APEgenspec;
  * QGModel (qg_model.py:213-263): the derived fields p, xi, Jptpc, Jpxi, Jq,
    Sph and APEflux, KEflux, APEgen;
  * LayeredModel (layered_model.py:242-301): p, xi, Jpxi, Jq, Sph, Sp, JSp,
    phn and KEspec_modal, PEspec_modal, APEspec, KEflux_div, APEflux_div;
  * the accumulation rule of _calc_diagnostics / _increment_diagnostics
    (model.py:577-580, 804-822);
  * the vertical modes (model.py:284-313) and the diagnostic constants of each
    model's constructor (qg_model.py:114-138, layered_model.py:120-209,
    bt_model.py:42-43, sqg_model.py:35-36).

Pinned in tests/test_oracle.py against the time means written from the
reference build (tests/golden/make_golden.py, `diagnostics_*` fixtures) and
against the reference's own identities (test_diagnostics.py:29-160).
"""
import numpy as np

from .qg_oracle import OracleBT, OracleLayered, OracleQG, OracleSQG

CORE = ["Ensspec", "KEspec", "EKEdiss", "KEfrictionspec", "EKE", "Dissspec", "ENSDissspec",
        "paramspec", "ENSparamspec"]
QGDIAG = ["entspec", "paramspec_APEflux", "paramspec_KEflux", "ENSflux", "ENSgenspec",
          "ENSfrictionspec", "APEgenspec"]
QG_ONLY = ["APEflux", "KEflux", "APEgen"]
LAYERED_ONLY = ["KEspec_modal", "PEspec_modal", "APEspec", "KEflux_div", "APEflux_div"]


def names(m):
    """Diagnostics the reference registers for this model class."""
    if isinstance(m, OracleQG):
        return CORE + QGDIAG + QG_ONLY
    if isinstance(m, OracleLayered):
        return CORE + QGDIAG + LAYERED_ONLY
    return list(CORE)


def attach_constants(m):
    """Diagnostic constants of the reference constructors."""
    if isinstance(m, OracleQG):
        m.U = m.U1 - m.U2
        m.Vbg = np.zeros_like(m.Ubg)
        m.del1 = m.delta / (m.delta + 1.)
        m.del2 = (m.delta + 1.) ** -1
        m.S = np.array([[-m.F1, m.F1], [m.F2, -m.F2]]).astype(np.float64)
        m.ikQy = np.vstack([(m.Qy[0] * 1j * m.k)[np.newaxis], (m.Qy[1] * 1j * m.k)[np.newaxis]])
        m.ilQx = 0.
    elif isinstance(m, OracleLayered):
        m.Ubg = m._Ubg
        m.Qx = np.dot(m.S, m.Vbg)
        m.ikQy = m.Qy[:, np.newaxis, np.newaxis] * 1j * m.k
        m.ilQx = m.Qx[:, np.newaxis, np.newaxis] * 1j * m.l
        if m.nz == 2 and m.rd and m.delta:
            m.del1 = m.delta / (m.delta + 1.)
            m.del2 = (m.delta + 1.) ** -1
        if m.nz == 2:
            m.f2gpi = np.array(m.rd ** -2 * (m.Hi[0] * m.Hi[1]) / m.H)[np.newaxis, np.newaxis]
        else:
            m.f2gpi = (m.f2 / m.gpi)[:, np.newaxis, np.newaxis]
        vertical_modes(m)
    elif isinstance(m, (OracleBT, OracleSQG)):
        m.Hi = np.array(m.H)[np.newaxis, ...]
    return m


def vertical_modes(m):
    """model.py:284-305"""
    evals, evecs = np.linalg.eig(-m.S)
    asort = evals.argsort()
    m.kdi2 = evals[asort]
    m.pmodes = evecs[:, asort]
    Ai = (m.H / (m.Hi[:, np.newaxis] * (m.pmodes ** 2)).sum(axis=0)) ** 0.5
    m.pmodes = Ai[np.newaxis, :] * m.pmodes


def modal_projection(m, p):
    """model.py:307-317 (forward): pmodes^-1 p at every wavenumber.  The
    reference's broadcast form needs numpy 1.x semantics of np.linalg.solve;
    this is the same per-wavenumber solve with a matrix right-hand side."""
    return np.linalg.solve(m.pmodes, p.reshape(m.nz, -1)).reshape(p.shape)


def advect(m, q, u=None, v=None):
    """model.py:538-558"""
    u = m.u if u is None else u
    v = m.v if v is None else v
    uq = u * q
    vq = v * q
    is_2d = uq.ndim == 2
    if is_2d:
        uq = np.tile(uq[np.newaxis, :, :], (m.nz, 1, 1))
        vq = np.tile(vq[np.newaxis, :, :], (m.nz, 1, 1))
    ik = m._ik[np.newaxis, np.newaxis, :]
    il = m._il[np.newaxis, :, np.newaxis]
    tend = ik * m.fft(uq) + il * m.fft(vq)
    return tend[0] if is_2d else tend


def derived_fields(m):
    """qg_model.py:213-224 and layered_model.py:242-254"""
    m.p = m.ifft(m.ph)
    m.xi = m.ifft(-m.wv2 * m.ph)
    m.Jpxi = advect(m, m.xi, m.u, m.v)
    m.Jq = advect(m, m.q, m.u, m.v)
    m.Sph = np.einsum("ij,jkl->ikl", m.S, m.ph) if hasattr(m, "S") else None
    if isinstance(m, OracleQG):
        m.Jptpc = -advect(m, m.p[0] - m.p[1],
                          m.del1 * m.u[0] + m.del2 * m.u[1],
                          m.del1 * m.v[0] + m.del2 * m.v[1])
    if isinstance(m, OracleLayered):
        m.Sp = m.ifft(m.Sph)
        m.JSp = advect(m, m.Sp, m.u, m.v)
        m.phn = modal_projection(m, m.ph)


def _dissipation_spectrum(m):
    """model.py:680-704"""
    spectrum = np.zeros_like(m.qh)
    ones = np.ones_like(m.filtr)
    if m.ablevel == 0:
        dt1, dt2, dt3 = m.dt, 0.0, 0.0
    elif m.ablevel == 1:
        dt1, dt2, dt3 = 1.5 * m.dt, -0.5 * m.dt, 0.0
    else:
        dt1, dt2, dt3 = 23. / 12. * m.dt, -16. / 12. * m.dt, 5. / 12. * m.dt
    for k in range(m.nz):
        spectrum[k] = (m.filtr - ones) * (m.qh[k] + dt1 * m.dqhdt[k] + dt2 * m.dqhdt_p[k]
                                          + dt3 * m.dqhdt_pp[k])
    return spectrum


def _param_contribution(m):
    """model.py:738-746"""
    dqh = np.zeros_like(m.qh)
    if m.uv_parameterization is not None:
        ik = np.asarray(m._ik).reshape((1, -1)).repeat(m.wv2.shape[0], axis=0)
        il = np.asarray(m._il).reshape((-1, 1)).repeat(m.wv2.shape[-1], axis=-1)
        dqh += -il * m.duh + ik * m.dvh
    if m.q_parameterization is not None:
        dqh += m.dqh
    return dqh


def _paramspec_contribution(m, term):
    """qg_diagnostics.py:72-74"""
    hr = (m.Hi / m.H)[:, np.newaxis, np.newaxis]
    return np.real(hr * m.ph.conj() * term).sum(axis=0) / m.M ** 2


def evaluate(m, name):
    """Instantaneous value of one diagnostic (derived_fields(m) must be current)."""
    M2 = m.M ** 2
    Hi3 = m.Hi[:, np.newaxis, np.newaxis]
    if name == "Ensspec":
        return np.abs(m.qh) ** 2 / M2
    if name == "KEspec":
        return m.wv2 * np.abs(m.ph) ** 2 / M2
    if name == "EKEdiss":
        return m.Hi[-1] / m.H * m.rek * (m.v[-1] ** 2 + m.u[-1] ** 2).mean()
    if name == "KEfrictionspec":
        return -m.rek * m.Hi[-1] / m.H * m.wv2 * np.abs(m.ph[-1]) ** 2 / M2
    if name == "EKE":
        return 0.5 * (m.v ** 2 + m.u ** 2).mean(axis=-1).mean(axis=-1)
    if name == "Dissspec":
        return -np.tensordot(m.Hi, np.conj(m.ph) * _dissipation_spectrum(m),
                             axes=(0, 0)).real / m.H / m.dt / M2
    if name == "ENSDissspec":
        return np.tensordot(m.Hi, np.conj(m.qh) * _dissipation_spectrum(m),
                            axes=(0, 0)).real / m.H / m.dt / M2
    if name == "paramspec":
        dqh = _param_contribution(m)
        return -np.real((Hi3 / m.H * np.conj(m.ph) * dqh).sum(axis=0)) / M2
    if name == "ENSparamspec":
        dqh = _param_contribution(m)
        return np.real((Hi3 / m.H * np.conj(m.qh) * dqh).sum(axis=0)) / M2
    if name == "entspec":
        return np.abs((Hi3 * m.qh).sum(axis=0) / m.H) ** 2 / M2
    if name == "paramspec_APEflux":
        return -_paramspec_contribution(
            m, np.einsum("ij, jk..., k... -> i...", m.S, m.a.real, _param_contribution(m)))
    if name == "paramspec_KEflux":
        return m.wv2 * _paramspec_contribution(
            m, np.einsum("ij..., j... -> i...", m.a.real, _param_contribution(m)))
    if name == "ENSflux":
        return (-Hi3 * (m.qh.conj() * m.Jq).real).sum(axis=0) / m.H / M2
    if name == "ENSgenspec":
        return (Hi3 * ((m.ilQx - m.ikQy) * m.Sph.conj() * m.ph).real).sum(axis=0) / m.H / M2
    if name == "ENSfrictionspec":
        return m.rek * m.Hi[-1] / m.H * m.wv2 * (m.qh[-1].conj() * m.ph[-1]).real / M2
    if name == "APEgenspec":
        return (Hi3 * (m.Ubg[:, np.newaxis, np.newaxis] * m.k
                       + m.Vbg[:, np.newaxis, np.newaxis] * m.l)
                * (1j * m.ph.conj() * m.Sph).real).sum(axis=0) / m.H / M2
    if name == "APEflux":
        return (m.rd ** -2 * m.del1 * m.del2
                * np.real((m.ph[0] - m.ph[1]) * np.conj(m.Jptpc)) / M2)
    if name == "KEflux":
        return (np.real(m.del1 * m.ph[0] * np.conj(m.Jpxi[0]))
                + np.real(m.del2 * m.ph[1] * np.conj(m.Jpxi[1]))) / M2
    if name == "APEgen":
        return (m.U * m.rd ** -2 * m.del1 * m.del2
                * np.real((1j * m.k * (m.del1 * m.ph[0] + m.del2 * m.ph[1])
                           * np.conj(m.ph[0] - m.ph[1])).sum()
                          + (1j * m.k[:, 1:-2] * (m.del1 * m.ph[0, :, 1:-2]
                                                  + m.del2 * m.ph[1, :, 1:-2])
                             * np.conj(m.ph[0, :, 1:-2] - m.ph[1, :, 1:-2])).sum()) / M2)
    if name == "KEspec_modal":
        return m.wv2 * (np.abs(m.phn) ** 2) / M2
    if name == "PEspec_modal":
        return m.kdi2[1:, np.newaxis, np.newaxis] * (np.abs(m.phn[1:, :, :]) ** 2) / M2
    if name == "APEspec":
        return (m.f2gpi * np.abs(m.ph[:-1] - m.ph[1:]) ** 2).sum(axis=0) / m.H / M2
    if name == "KEflux_div":
        return (Hi3 * (m.ph.conj() * m.Jpxi).real).sum(axis=0) / m.H / M2
    if name == "APEflux_div":
        return (Hi3 * (m.ph.conj() * m.JSp).real).sum(axis=0) / m.H / M2
    raise KeyError(name)


class Accumulator:
    """The reference's registry arithmetic: value = sum of the instantaneous
    values at the accumulation steps, get(name) = value / count."""

    def __init__(self, model, tavestart=0., taveint=None, active=None):
        self.m = attach_constants(model)
        self.tavestart = tavestart
        taveint = model.dt if taveint is None else taveint
        self.taveints = np.ceil(taveint / model.dt)
        self.active = list(names(model) if active is None else active)
        self.value = {}
        self.count = 0

    def due(self):
        """model.py:577-580; BTModel and SQGModel skip the tavestart test
        (bt_model.py:106-109, sqg_model.py:75-78)"""
        m = self.m
        if isinstance(m, (OracleBT, OracleSQG)):
            return (m.t >= m.dt) and (m.tc % self.taveints == 0)
        return (m.t >= m.dt) and (m.t >= self.tavestart) and (m.tc % self.taveints == 0)

    def increment(self):
        derived_fields(self.m)
        for n in self.active:
            v = np.array(evaluate(self.m, n), dtype=np.float64)
            self.value[n] = v if n not in self.value else self.value[n] + v
        self.count += 1

    def get(self, name):
        return self.value[name] / self.count


def step_with_diagnostics(m, acc):
    """_step_forward (model.py:406-436) with _calc_diagnostics in its place."""
    m._invert()
    m._do_advection()
    m._do_friction()
    if m.uv_parameterization is not None:
        m._do_uv_subgrid_parameterization()
    if m.q_parameterization is not None:
        m._do_q_subgrid_parameterization()
    if acc.due():
        acc.increment()
    m._forward_timestep()
