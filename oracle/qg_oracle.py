"""CPU oracle: a NumPy FP64 restatement of pyqg's time-stepping kernel.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this module;
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, and
only as the checker (or as the timed CPU baseline), never as the thing measured.

It restates, operation for operation, the numpy.fft fallback build of
pyqg/kernel.pyx (`PYQG_USE_PYFFTW == 0`, kernel.pyx:245-267): the same FFT
conventions (rfftn / irfftn over axes (-2,-1); irfftn discards the imaginary
parts of the x-Nyquist and x-DC columns after the y transform), the same order
of floating-point operations in every pointwise loop, the same AB1/AB2/AB3
start-up and filter placement.  The model-level constant builders restate
pyqg/model.py, qg_model.py, layered_model.py, bt_model.py and sqg_model.py.

How each part is pinned (tests/test_oracle.py):
  * FFTs: numpy.fft is the reference's own transform in this build; bin
    amplitudes (test_model.py:16-65), Parseval (test_fft2.py:8-40), and a
    brute-force O(N^2) DFT at nx=8.
  * _invert / _do_advection / _do_friction / _forward_timestep: the reference's
    closed-form tests (test_model.py:67-340, test_advection.py) restated, plus
    the golden trajectories written from the reference build
    (tests/golden/make_golden.py) at 1e-13.
  * Composition over long runs: the regression values held by the reference
    (test_reference_solns.py) and the same runs of the reference build.
  * Constants (a, filtr, Qy, kk, ll, S): golden vectors; N-layer `a` also
    against a per-wavenumber numpy.linalg.inv and the 3x3 cofactor check of
    test_layered.py:60-84.
"""
import numpy as np
from numpy import pi

rfftn = np.fft.rfftn
irfftn = np.fft.irfftn
AX = (-2, -1)


class OracleKernel:
    """Restates `cdef class PseudoSpectralKernel` (pyqg/kernel.pyx:32-637)."""

    def __init__(self, nz, ny, nx, fftw_num_threads=1, has_q_param=0, has_uv_param=0):
        # kernel.pyx:120-206
        self.nz, self.ny, self.nx = nz, ny, nx
        self.nl = ny
        self.nk = int(nx / 2 + 1)
        self._a = np.zeros((nz, nz, self.nl, self.nk), np.complex128)
        self._kk = np.zeros(self.nk)
        self._ik = np.zeros(self.nk, np.complex128)
        self._ll = np.zeros(self.nl)
        self._il = np.zeros(self.nl, np.complex128)
        self._k2l2 = np.zeros((self.nl, self.nk))
        R = lambda: np.zeros((nz, ny, nx))
        C = lambda: np.zeros((nz, self.nl, self.nk), np.complex128)
        self._q, self._qh, self._ph = R(), C(), C()
        self._u, self._uh, self._v, self._vh = R(), C(), R(), C()
        self._uq, self._uqh, self._vq, self._vqh = R(), C(), R(), C()
        if has_uv_param:
            self._du, self._dv, self._duh, self._dvh = R(), R(), C(), C()
        if has_q_param:
            self._dq, self._dqh = R(), C()
        self._dt = 0.0
        self.t = 0.0
        self.tc = 0
        self.ablevel = 0
        self.rek = 0.0
        self._dqhdt, self._dqhdt_p, self._dqhdt_pp = C(), C(), C()
        if not hasattr(self, "_Ubg"):
            self._Ubg = np.zeros(nz)
        if not hasattr(self, "_Qy"):
            self._Qy = np.zeros(nz)
        self._ikQy = np.zeros((nz, self.nk), np.complex128)
        self.filtr = np.ones((self.nl, self.nk))

    # -- generic transforms (kernel.pyx:291-309): copy in, transform, return a copy
    def fft(self, v):
        return rfftn(np.array(v, dtype=np.float64), axes=AX)

    def ifft(self, v):
        return irfftn(np.array(v, dtype=np.complex128), axes=AX)

    # -- kernel.pyx:312-355
    def _invert(self):
        ph = np.zeros_like(self._qh)
        for k2 in range(self.nz):
            for k1 in range(self.nz):
                ph[k2] = ph[k2] + self._a[k2, k1] * self._qh[k1]
        self._ph = ph
        self._uh = -self._il[np.newaxis, :, np.newaxis] * ph
        self._vh = self._ik[np.newaxis, np.newaxis, :] * ph
        self._u = irfftn(self._uh, axes=AX)
        self._v = irfftn(self._vh, axes=AX)

    # -- kernel.pyx:357-395 (no Vbg term, no ilQx term)
    def _do_advection(self):
        self._uq = (self._u + self._Ubg[:, np.newaxis, np.newaxis]) * self._q
        self._vq = self._v * self._q
        self._uqh = rfftn(self._uq, axes=AX)
        self._vqh = rfftn(self._vq, axes=AX)
        self._dqhdt = -(self._ik[np.newaxis, np.newaxis, :] * self._uqh
                        + self._il[np.newaxis, :, np.newaxis] * self._vqh
                        + self._ikQy[:, np.newaxis, :] * self._ph)

    # -- kernel.pyx:397-422
    def _do_uv_subgrid_parameterization(self):
        du, dv = self.uv_parameterization(self)
        self._du[:] = du
        self._dv[:] = dv
        self._duh = rfftn(self._du, axes=AX)
        self._dvh = rfftn(self._dv, axes=AX)
        self._dqhdt = self._dqhdt + (-self._il[np.newaxis, :, np.newaxis] * self._duh
                                     + self._ik[np.newaxis, np.newaxis, :] * self._dvh)

    # -- kernel.pyx:424-442
    def _do_q_subgrid_parameterization(self):
        dq = self.q_parameterization(self)
        self._dq[:] = dq
        self._dqh = rfftn(self._dq, axes=AX)
        self._dqhdt = self._dqhdt + self._dqh

    # -- kernel.pyx:444-461: bottom layer only, skipped when rek == 0
    def _do_friction(self):
        if self.rek:
            k = self.nz - 1
            self._dqhdt[k] = self._dqhdt[k] + (self.rek * self._k2l2) * self._ph[k]

    # -- kernel.pyx:463-526
    def _forward_timestep(self):
        if self.ablevel == 0:
            dt1, dt2, dt3 = self._dt, 0.0, 0.0
            self.ablevel = 1
        elif self.ablevel == 1:
            dt1, dt2, dt3 = 1.5 * self._dt, -0.5 * self._dt, 0.0
            self.ablevel = 2
        else:
            dt1, dt2, dt3 = 23. / 12. * self._dt, -16. / 12. * self._dt, 5. / 12. * self._dt
        qh_new = self.filtr * (self._qh + dt1 * self._dqhdt + dt2 * self._dqhdt_p
                               + dt3 * self._dqhdt_pp)
        self._qh = qh_new
        self._dqhdt_pp = self._dqhdt_p.copy()
        self._dqhdt_p = self._dqhdt.copy()
        self._q = irfftn(qh_new, axes=AX)
        self.tc += 1
        self.t += self._dt

    # -- properties (kernel.pyx:528-637)
    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, dt):
        self._dt = dt
        self.ablevel = 0

    @property
    def kk(self):
        return self._kk

    @kk.setter
    def kk(self, k):
        self._kk = np.asarray(k, dtype=np.float64)
        self._ik = 1j * self._kk
        self._k2l2 = self._kk[np.newaxis, :] ** 2 + self._ll[:, np.newaxis] ** 2

    @property
    def ll(self):
        return self._ll

    @ll.setter
    def ll(self, l):
        self._ll = np.asarray(l, dtype=np.float64)
        self._il = 1j * self._ll
        self._k2l2 = self._kk[np.newaxis, :] ** 2 + self._ll[:, np.newaxis] ** 2

    @property
    def a(self):
        return self._a

    @a.setter
    def a(self, b):
        self._a[:] = np.asarray(b, dtype=np.float64).astype(np.complex128)

    @property
    def Ubg(self):
        return self._Ubg

    @Ubg.setter
    def Ubg(self, U):
        self._Ubg = np.asarray(U, dtype=np.float64)

    @property
    def Qy(self):
        return self._Qy

    @Qy.setter
    def Qy(self, Qy):
        self._Qy = np.asarray(Qy, dtype=np.float64)
        self._ikQy = 1j * (self._kk[np.newaxis, :] * self._Qy[:, np.newaxis])

    @property
    def q(self):
        return self._q

    @q.setter
    def q(self, b):
        self._q[:] = b
        self._qh = rfftn(self._q, axes=AX)

    @property
    def qh(self):
        return self._qh

    @qh.setter
    def qh(self, b):
        self._qh = np.array(b, dtype=np.complex128)
        self._q = irfftn(self._qh, axes=AX)


def _ro(name):
    return property(lambda self: getattr(self, name))


for _f in ["ph", "u", "v", "uh", "vh", "uq", "vq", "uqh", "vqh", "dqhdt", "dqhdt_p",
           "dqhdt_pp", "dq", "dqh", "du", "dv", "duh", "dvh"]:
    setattr(OracleKernel, _f, _ro("_" + _f))
OracleKernel.ufull = property(lambda self: self._u + self._Ubg[:, np.newaxis, np.newaxis])
OracleKernel.vfull = property(lambda self: self._v)


class OracleModel(OracleKernel):
    """Restates Model.__init__ (pyqg/model.py:101-258), the grid (446-482), the
    filter (496-503) and the step order of _step_forward (406-436).  Diagnostics
    and logging are not restated (they do not touch the state)."""

    def __init__(self, nz=1, nx=64, ny=None, L=1e6, W=None, dt=7200., tmax=1576800000.,
                 rek=5.787e-7, filterfac=23.6, f=None, g=9.81, q_parameterization=None,
                 uv_parameterization=None, **unused):
        ny = nx if ny is None else ny
        W = L if W is None else W
        OracleKernel.__init__(self, nz, ny, nx, 1, int(q_parameterization is not None),
                              int(uv_parameterization is not None))
        self.L, self.W = L, W
        self.dt = dt
        self.tmax = tmax
        self.rek = rek
        self.filterfac = filterfac
        self.g = g
        if f:
            self.f = f
            self.f2 = f ** 2
        self.q_parameterization = q_parameterization
        self.uv_parameterization = uv_parameterization
        self._init_grid()
        self._init_background()
        self._init_filter()
        self._init_inversion()

    def _init_grid(self):
        self.x, self.y = np.meshgrid(np.arange(0.5, self.nx, 1.) / self.nx * self.L,
                                     np.arange(0.5, self.ny, 1.) / self.ny * self.W)
        self.dk = 2. * pi / self.L
        self.dl = 2. * pi / self.W
        # model.py:461-462: built from nx (only nx == ny is meaningful)
        self.ll = self.dl * np.append(np.arange(0., self.nx / 2), np.arange(-self.nx / 2, 0.))
        self.kk = self.dk * np.arange(0., self.nk)
        self.k, self.l = np.meshgrid(self.kk, self.ll)
        self.dx = self.L / self.nx
        self.dy = self.W / self.ny
        self.M = self.nx * self.ny
        self.wv2 = self.k ** 2 + self.l ** 2
        self.wv = np.sqrt(self.wv2)
        nz2 = self.wv2 != 0.
        self.wv2i = np.zeros_like(self.wv2)
        self.wv2i[nz2] = self.wv2[nz2] ** -1

    def _init_filter(self):
        cphi = 0.65 * pi
        wvx = np.sqrt((self.k * self.dx) ** 2. + (self.l * self.dy) ** 2.)
        filtr = np.exp(-self.filterfac * (wvx - cphi) ** 4.)
        filtr[wvx <= cphi] = 1.
        self.filtr = filtr

    def _step_forward(self):
        self._invert()
        self._do_advection()
        self._do_friction()
        if self.uv_parameterization is not None:
            self._do_uv_subgrid_parameterization()
        if self.q_parameterization is not None:
            self._do_q_subgrid_parameterization()
        self._forward_timestep()

    def run(self):
        while self.t < self.tmax:
            self._step_forward()

    def spec_var(self, ph):
        """model.py:832-838"""
        var_dens = 2. * np.abs(ph) ** 2 / self.M ** 2
        var_dens[..., 0] = var_dens[..., 0] / 2.
        var_dens[..., -1] = var_dens[..., -1] / 2.
        return var_dens.sum()


class OracleQG(OracleModel):
    """qg_model.py:60-161"""

    def __init__(self, beta=1.5e-11, rd=15000.0, delta=0.25, H1=500, U1=0.025, U2=0.0, **kw):
        self.beta, self.rd, self.delta = beta, rd, delta
        self.Hi = np.array([H1, H1 / delta])
        self.U1, self.U2 = U1, U2
        super().__init__(nz=2, **kw)

    def _init_background(self):
        self.H = self.Hi.sum()
        self.Ubg = np.array([self.U1, self.U2])
        self.F1 = self.rd ** -2 / (1. + self.delta)
        self.F2 = self.delta * self.F1
        self.Qy = np.array([self.beta + self.F1 * (self.U1 - self.U2),
                            self.beta - self.F2 * (self.U1 - self.U2)])

    def _init_inversion(self):
        a = np.ma.zeros((2, 2, self.nl, self.nk), np.dtype('float64'))
        det_inv = np.ma.masked_equal(self.wv2 * (self.wv2 + self.F1 + self.F2), 0.) ** -1
        a[0, 0] = -(self.wv2 + self.F2) * det_inv
        a[0, 1] = -self.F1 * det_inv
        a[1, 0] = -self.F2 * det_inv
        a[1, 1] = -(self.wv2 + self.F1) * det_inv
        self.a = np.ma.masked_invalid(a).filled(0.)

    def set_q1q2(self, q1, q2):
        self.q = np.vstack([q1[np.newaxis], q2[np.newaxis]])

    def _calc_ke(self):
        ke = 0.
        for j in range(self.nz):
            ke += .5 * self.Hi[j] * self.spec_var(self.wv * self._ph[j])
        return ke.sum() / self.H


class OracleLayered(OracleModel):
    """layered_model.py:72-229"""

    def __init__(self, beta=1.5e-11, nz=3, rd=15000.0, f=0.0001236812857687059, H=None,
                 U=None, V=None, rho=None, delta=None, **kw):
        self.beta, self.rd, self.delta = beta, rd, delta
        if U is None:
            U = (np.arange(nz) * 0.025)[::-1]
        if V is None:
            V = np.zeros(nz)
        if H is None:
            H = [500] + [1750 for _ in range(nz - 1)]
        if rho is None:
            rho = np.arange(nz) * 0.3 + 1025
        self._Ubg = np.array(U, dtype=np.float64)
        self.Vbg = np.array(V)
        self.Hi = np.array(H)
        self.rhoi = np.array(rho)
        super().__init__(nz=nz, f=f, **kw)

    def _init_background(self):
        self.H = self.Hi.sum()
        nz = self.nz
        S = np.zeros((nz, nz))
        if nz == 2 and self.rd and self.delta:
            F1 = self.rd ** -2 / (1. + self.delta)
            F2 = self.delta * F1
            S[0, 0], S[0, 1] = -F1, F1
            S[1, 0], S[1, 1] = F2, -F2
        else:
            self.gpi = self.g * (self.rhoi[1:] - self.rhoi[:-1]) / self.rhoi[:-1]
            assert np.all(self.gpi > 0.), "Buoyancy jump has negative sign!"
            for i in range(nz):
                if i == 0:
                    S[i, i] = -self.f2 / self.Hi[i] / self.gpi[i]
                    S[i, i + 1] = self.f2 / self.Hi[i] / self.gpi[i]
                elif i == nz - 1:
                    S[i, i] = -self.f2 / self.Hi[i] / self.gpi[i - 1]
                    S[i, i - 1] = self.f2 / self.Hi[i] / self.gpi[i - 1]
                else:
                    S[i, i - 1] = self.f2 / self.Hi[i] / self.gpi[i - 1]
                    S[i, i] = -(self.f2 / self.Hi[i] / self.gpi[i]
                                + self.f2 / self.Hi[i] / self.gpi[i - 1])
                    S[i, i + 1] = self.f2 / self.Hi[i] / self.gpi[i]
        self.S = S
        self.Qy = self.beta - np.dot(S, self._Ubg)

    def _init_inversion(self):
        if self.nz == 2:
            S = self.S
            a = np.ma.zeros((2, 2, self.nl, self.nk), np.dtype('float64'))
            det_inv = np.ma.masked_equal((S[0, 0] - self.wv2) * (S[1, 1] - self.wv2)
                                         - S[0, 1] * S[1, 0], 0.) ** -1
            a[0, 0] = (S[1, 1] - self.wv2) * det_inv
            a[0, 1] = -S[0, 1] * det_inv
            a[1, 0] = -S[1, 0] * det_inv
            a[1, 1] = (S[0, 0] - self.wv2) * det_inv
        else:
            I = np.eye(self.nz)[:, :, np.newaxis, np.newaxis]
            M = self.S[:, :, np.newaxis, np.newaxis] - I * self.wv2
            M[:, :, 0, 0] = np.nan
            a = np.linalg.inv(M.T).T
        self.a = np.ma.masked_invalid(a).filled(0.)

    def _calc_ke(self):
        ke = 0.
        for j in range(self.nz):
            ke += .5 * self.Hi[j] * self.spec_var(self.wv * self._ph[j])
        return ke.sum() / self.H


class OracleBT(OracleModel):
    """bt_model.py:26-79"""

    def __init__(self, beta=0., rd=None, H=1., U=0., **kw):
        self.beta, self.rd, self.H, self.U = beta, rd, H, U
        self.kd2 = rd ** -2 if rd else 0.
        super().__init__(nz=1, **kw)

    def _init_background(self):
        self.Qy = np.asarray(self.beta, dtype=np.float64)[np.newaxis]
        self.Ubg = np.asarray(self.U, dtype=np.float64)[np.newaxis]

    def _init_inversion(self):
        if self.rd:
            self.a = -(1. / (self.wv2 + self.kd2))[np.newaxis, np.newaxis]
        else:
            self.a = -self.wv2i[np.newaxis, np.newaxis]

    def _calc_ke(self):
        return (.5 * self.spec_var(self.wv * self._ph)).sum()


class OracleSQG(OracleModel):
    """sqg_model.py:9-66"""

    def __init__(self, beta=0., Nb=1., f_0=1., H=1., U=0., **kw):
        self.beta, self.Nb, self.f_0, self.H, self.U = beta, Nb, f_0, H, U
        super().__init__(nz=1, **kw)

    def _init_background(self):
        self.Qy = np.asarray(self.beta, dtype=np.float64)[np.newaxis]
        self.Ubg = np.asarray(self.U, dtype=np.float64)[np.newaxis]

    def _init_inversion(self):
        self.a = np.asarray(self.f_0 / self.Nb * np.sqrt(self.wv2i))[np.newaxis, np.newaxis]

    def _calc_ke(self):
        return (.5 * self.spec_var(self.wv * self._ph)).sum()


def ab_coefficients(ablevel, dt):
    """The AB1/AB2/AB3 weights of kernel.pyx:482-500 for a given ablevel."""
    if ablevel == 0:
        return dt, 0.0, 0.0
    if ablevel == 1:
        return 1.5 * dt, -0.5 * dt, 0.0
    return 23. / 12. * dt, -16. / 12. * dt, 5. / 12. * dt
