"""Model driver and the four physics front-ends on top of the HIP kernel.

The classes keep the constructor signatures, initialisation order and stepping
order of pyqg's Model / QGModel / LayeredModel / BTModel / SQGModel
(pyqg/model.py:101-258, 406-436; qg_model.py; layered_model.py; bt_model.py;
sqg_model.py).  They only build the constants the kernel consumes (wavenumbers,
filter, background, inversion matrix) and orchestrate the kernel's phases;
all per-step arithmetic happens in libqgk, including the time-averaged
diagnostics (pyqg_amd.diagnostics).  Linear stability analysis is not part of
this package; xarray output (Model.to_dataset) is in pyqg_amd.dataset.
"""
import logging
import warnings

import numpy as np
from numpy import pi

from . import _lib
from . import constants as C
from . import diagnostics as Dg
from .diagnostics import DiagnosticsMixin
from .kernel import PseudoSpectralKernel


class Model(DiagnosticsMixin, PseudoSpectralKernel):
    """Generic pseudo-spectral model (pyqg/model.py:22-258)."""

    def __init__(self, nz=1, nx=64, ny=None, L=1e6, W=None, dt=7200., twrite=1000.,
                 tmax=1576800000., tavestart=315360000., taveint=86400., useAB2=False,
                 rek=5.787e-7, filterfac=23.6, f=None, g=9.81, q_parameterization=None,
                 uv_parameterization=None, parameterization=None, diagnostics_list='all',
                 ntd=1, log_level=1, logfile=None):
        if ny is None:
            ny = nx
        if W is None:
            W = L
        # infer the type of an explicit parameterization object (model.py:201-212)
        if parameterization is not None:
            ptype = getattr(parameterization, "parameterization_type", None)
            if ptype == 'uv_parameterization':
                assert uv_parameterization is None
                uv_parameterization = parameterization
            elif ptype == 'q_parameterization':
                assert q_parameterization is None
                q_parameterization = parameterization
            else:
                raise ValueError(f"unknown parameterization type {ptype}")

        PseudoSpectralKernel.__init__(self, nz, ny, nx, ntd,
                                      has_q_param=int(q_parameterization is not None),
                                      has_uv_param=int(uv_parameterization is not None))
        self.L, self.W = L, W
        self.dt = dt
        self.twrite = twrite
        self.tmax = tmax
        self.tavestart = tavestart
        self.taveint = taveint
        self.logfile = logfile
        self.log_level = log_level
        self.useAB2 = useAB2
        self.ntd = ntd
        self.rek = rek
        self.filterfac = filterfac
        self.g = g
        if f:
            self.f = f
            self.f2 = f ** 2
        self.q_parameterization = q_parameterization
        self.uv_parameterization = uv_parameterization
        self.diagnostics_list = diagnostics_list

        self._initialize_logger()
        self._initialize_grid()
        self._initialize_background()
        self._initialize_forcing()
        self._initialize_filter()
        self._initialize_time()
        self._initialize_inversion_matrix()
        self._initialize_diagnostics(diagnostics_list)

    # ------------------------------------------------------------ running
    def run_with_snapshots(self, tsnapstart=0., tsnapint=432000.):
        """Yield model time every tsnapint once t >= tsnapstart (model.py:260-278).

        Between yields the steps run as in run(): fused stretches, with the
        step before each yield, status report or diagnostic accumulation run
        phase by phase, so the fields a caller reads at a yield are the
        per-method path's."""
        tsnapints = np.ceil(tsnapint / self.dt)
        if self.q_parameterization is not None or self.uv_parameterization is not None:
            while self.t < self.tmax:
                self._step_forward()
                if self.t >= tsnapstart and (self.tc % tsnapints) == 0:
                    yield self.t
            return
        while self.t < self.tmax:
            # steps up to and including the next one that yields (or the end)
            t, tc, n = self.t, self.tc, 0
            while t < self.tmax:
                t += self.dt
                tc += 1
                n += 1
                if t >= tsnapstart and (tc % tsnapints) == 0:
                    break
            self._advance(n)
            if self.t >= tsnapstart and (self.tc % tsnapints) == 0:
                yield self.t

    def run(self):
        """Run until t >= tmax (model.py:280-283).

        Without parameterizations the steps between two status reports or
        diagnostic accumulations run in the fused two-kernel form; the step
        that a status report observes runs phase by phase, so KE / CFL see the
        same fields as the reference's, and a step that accumulates
        diagnostics runs as _step_forward.  With built-in closures (evaluated
        on the device) those stretches replay a HIP graph of the phase-by-phase
        step; other parameterizations step one _step_forward at a time."""
        if ((self.q_parameterization is not None or self.uv_parameterization is not None)
                and not self._graphable()):
            while self.t < self.tmax:
                self._step_forward()
            return
        # how many steps until t >= tmax with the reference's accumulated time
        t, n = self.t, 0
        while t < self.tmax:
            t += self.dt
            n += 1
        self._advance(n)

    def _advance(self, n):
        """n steps: stretches ending at status reports and before diagnostic
        steps run fused (no parameterization) or as graph replays (built-in
        closures); the diagnostic steps run as _step_forward."""
        stretch = self._fused_steps
        if self.q_parameterization is not None or self.uv_parameterization is not None:
            if not self._graphable():
                for _ in range(n):
                    self._step_forward()
                return
            stretch = self._graph_steps
        while n > 0:
            chunk = n
            if self.log_level:
                to_status = int(self.twrite - (self.tc % self.twrite))
                chunk = min(n, max(to_status, 1))
            chunk = min(chunk, self._steps_before_diagnostics(chunk))
            if chunk == 0:
                self._step_forward()
                n -= 1
                continue
            stretch(chunk)
            self._print_status()
            n -= chunk

    def _graphable(self):
        """Every parameterization present is a built-in closure evaluated on the
        device, and the model adds no external forcing of its own."""
        for p in (self.q_parameterization, self.uv_parameterization):
            if p is not None and self._device_parameterization(p) is None:
                return False
        return type(self)._do_external_forcing is Model._do_external_forcing

    def _graph_steps(self, nsteps):
        """nsteps parameterized steps with no diagnostic or status step among
        them.  Once the Adams-Bashforth weights are constant the steps replay a
        HIP graph of two phase-by-phase steps (invert, advection, friction,
        closures written into du/dv or dq on the device, time step; two steps
        are one period of the tendency history's buffer swap), so every kernel and
        every field is the one _step_forward produces; the rest run as
        _step_forward.  The graph is kept while the closures, dt and the
        kernel's constants stay the same."""
        nsteps = int(nsteps)
        while nsteps > 0 and self._ablevel < 2:
            self._step_forward()      # AB start-up
            nsteps -= 1
        if nsteps >= 2:
            from .parameterizations import _capture_steps, _device_target
            L = _lib.lib
            b, dx, Hi = _device_target(self)
            qp, uvp = self.q_parameterization, self.uv_parameterization
            w, self._ablevel = C.ab_weights(self._ablevel, self._dt)
            key = (None if qp is None else qp._device_key(),
                   None if uvp is None else uvp._device_key(), tuple(w), dx.tobytes(),
                   Hi.tobytes())
            self._flush()
            ok = getattr(self, "_graph_key", None) == key
            if not ok or L.qgk_graph_replay(b.h, 0) != 0:
                # an ordinary step first: before a capture it makes the
                # allocations and uploads; for a kept graph it brings the
                # history buffers back to the phase the graph was captured at
                self._step_forward()
                nsteps -= 1
                ok = ok and L.qgk_graph_replay(b.h, 0) == 0
            if not ok and nsteps >= 2:
                self._graph_key = None
                _capture_steps(b, dx, Hi, qp, uvp, w)
                self._graph_key = key
                ok = True
            if ok and nsteps >= 2:
                reps = nsteps // 2
                _lib.check(L.qgk_graph_replay(b.h, reps))
                self._invalidate_all()
                for _ in range(2 * reps):
                    self._t += self._dt
                self._tc += 2 * reps
                nsteps -= 2 * reps
        for _ in range(nsteps):
            self._step_forward()

    def _steps_before_diagnostics(self, limit):
        """Steps (at most limit) that start before the next diagnostic step."""
        if not any(d['active'] for d in self.diagnostics.values()):
            return limit
        t, tc = self.t, self.tc
        for j in range(limit):
            if self._diagnostics_due(t, tc):
                return j
            t += self.dt
            tc += 1
        return limit

    def _step_forward(self):
        """One step in the reference's order (model.py:406-436)."""
        self._invert()
        self._do_advection()
        self._do_friction()
        self._do_external_forcing()
        if self.uv_parameterization is not None:
            self._do_uv_subgrid_parameterization()
        if self.q_parameterization is not None:
            self._do_q_subgrid_parameterization()
        self._calc_diagnostics()
        self._forward_timestep()
        self._print_status()

    def _do_external_forcing(self):
        pass

    def _print_status(self):
        """Log KE and CFL every twrite steps; assert CFL < 1 (model.py:564-574)."""
        if self.log_level and (self.tc % self.twrite) == 0:
            self.ke = self._calc_ke()
            self.cfl = self._calc_cfl()
            self.logger.info('Step: %i, Time: %3.2e, KE: %3.2e, CFL: %4.3f',
                             self.tc, self.t, self.ke, self.cfl)
            assert self.cfl < 1., self.logger.error('CFL condition violated')

    # ------------------------------------------------------------ set-up
    def _initialize_time(self):
        self.taveints = np.ceil(self.taveint / self.dt)

    def _initialize_grid(self):
        """Grids and wavenumbers (model.py:446-482)."""
        g = C.grid(self.nx, self.ny, self.L, self.W)
        self.x, self.y = g["x"], g["y"]
        self.dk, self.dl = g["dk"], g["dl"]
        self.ll = g["ll"]
        self.kk = g["kk"]
        self.k, self.l = g["k"], g["l"]
        self.ik = 1j * self.k
        self.il = 1j * self.l
        self.dx, self.dy, self.M = g["dx"], g["dy"], g["M"]
        self.wv2, self.wv, self.wv2i = g["wv2"], g["wv"], g["wv2i"]

    def _initialize_filter(self):
        """Exponential cutoff filter of Arbic & Flierl (2003) (model.py:496-503)."""
        self.filtr = C.spectral_filter(self.k, self.l, self.dx, self.dy, self.filterfac)

    def _filter(self, q):
        return self.filtr * q

    def _initialize_background(self):
        raise NotImplementedError('needs to be implemented by Model subclass')

    def _initialize_inversion_matrix(self):
        raise NotImplementedError('needs to be implemented by Model subclass')

    def _initialize_forcing(self):
        pass

    def _initialize_logger(self):
        self.logger = logging.getLogger(__name__)
        if self.logfile is not None:
            handler = logging.FileHandler(filename=self.logfile, mode='w')
        else:
            handler = logging.StreamHandler()
        handler.setFormatter(logging.Formatter('%(levelname)s: %(message)s'))
        if not self.logger.handlers:
            self.logger.addHandler(handler)
        self.logger.setLevel(self.log_level * 10)
        self.logger.propagate = False

    # ------------------------------------------------------------ helpers
    def spec_var(self, ph):
        """Variance of a field from its rfft2 coefficients (model.py:832-838)."""
        var_dens = 2. * np.abs(ph) ** 2 / self.M ** 2
        var_dens[..., 0] = var_dens[..., 0] / 2.
        var_dens[..., -1] = var_dens[..., -1] / 2.
        return var_dens.sum()

    def _advect(self, q, u=None, v=None):
        """Spectral advective tendency of real q, u, v (model.py:538-558)."""
        u = self.u if u is None else u
        v = self.v if v is None else v
        uq, vq = u * q, v * q
        is_2d = uq.ndim == 2
        if is_2d:
            uq = np.tile(uq[np.newaxis], (self.nz, 1, 1))
            vq = np.tile(vq[np.newaxis], (self.nz, 1, 1))
        tend = self.ik * self.fft(uq) + self.il * self.fft(vq)
        return tend[0] if is_2d else tend

    def set_q(self, q):
        warnings.warn("Method deprecated. Set model.q directly instead. ", DeprecationWarning)
        self.q = q

    def set_qh(self, qh):
        warnings.warn("Method deprecated. Set model.qh directly instead. ", DeprecationWarning)
        self.qh = qh

    def vertical_modes(self):
        """Deformation radii and pressure modes from the stretching matrix
        (model.py:286-305)."""
        self.kdi2, self.pmodes = C.vertical_modes(self.S, self.Hi)
        self.radii = np.zeros_like(self.kdi2)
        self.radii[0] = np.sqrt(self.g * self.H) / np.abs(self.f)
        self.radii[1:] = 1. / np.sqrt(self.kdi2[1:])

    def to_dataset(self):
        """The state, constants and accumulated diagnostics as an
        xarray.Dataset (model.py:840-848); ImportError without xarray.
        `pyqg_amd.dataset.dataset_parts(model)` gives the same contents as
        plain (dims, data, attrs) tuples."""
        from .dataset import model_to_dataset
        return model_to_dataset(self)

    @property
    def parameterization(self):
        if self.q_parameterization and self.uv_parameterization:
            warnings.warn("Model has multiple parameterizations, but only returning PV")
        return self.q_parameterization or self.uv_parameterization

    def _calc_cfl(self):
        return np.abs(np.hstack([self.u + self.Ubg[:, np.newaxis, np.newaxis], self.v])
                      ).max() * self.dt / self.dx

    def _calc_ke(self):
        ke = 0.
        for j in range(self.nz):
            ke += .5 * self.Hi[j] * self.spec_var(self.wv * self.ph[j])
        return ke.sum() / self.H


class QGModel(Model):
    """Two-layer QG model (pyqg/qg_model.py:17-209)."""

    def __init__(self, beta=1.5e-11, rd=15000.0, delta=0.25, H1=500, U1=0.025, U2=0.0, **kwargs):
        self.beta = beta
        self.rd = rd
        self.delta = delta
        self.Hi = np.array([H1, H1 / delta])
        self.U1 = U1
        self.U2 = U2
        super().__init__(nz=2, **kwargs)
        # initial PV anomalies (qg_model.py:103-106)
        self.set_q1q2(1e-7 * np.random.rand(self.ny, self.nx)
                      + 1e-6 * (np.ones((self.ny, 1)) * np.random.rand(1, self.nx)),
                      np.zeros_like(self.x))

    def _initialize_background(self):
        """qg_model.py:110-138"""
        self.H = self.Hi.sum()
        self.set_U1U2(self.U1, self.U2)
        self.U = self.U1 - self.U2
        self.Vbg = np.zeros_like(self.Ubg)
        self.F1, self.F2, Qy, _ = C.qg_background(self.beta, self.rd, self.delta, self.U1,
                                                  self.U2)
        self.Qy1, self.Qy2 = Qy
        self.Qy = Qy
        self.ikQy1 = self.Qy1 * 1j * self.k
        self.ikQy2 = self.Qy2 * 1j * self.k
        self.ikQy = np.vstack([self.ikQy1[np.newaxis], self.ikQy2[np.newaxis]])
        self.ilQx = 0.
        self.del1 = self.delta / (self.delta + 1.)
        self.del2 = (self.delta + 1.) ** -1

    @property
    def S(self):
        return np.array([[-self.F1, self.F1], [self.F2, -self.F2]]).astype(np.float64)

    def _initialize_inversion_matrix(self):
        """Closed-form 2x2 inverse of (S - kappa^2 I), zero at kappa = 0
        (qg_model.py:146-161)."""
        self.a = C.qg_inversion(self.wv2, self.F1, self.F2)

    def _initialize_model_diagnostics(self):
        """qg_diagnostics.py:6-70 and qg_model.py:226-263"""
        self._add_device_diagnostics(Dg.QGDIAG)
        self._add_device_diagnostics(Dg.QG_ONLY)

    def _calc_derived_fields(self):
        """Host fields of pyqg's diagnostic functions (qg_model.py:213-224), for
        diagnostics added with add_diagnostic; the built-in ones compute their
        own on the device."""
        self.p = self.ifft(self.ph)
        self.xi = self.ifft(-self.wv2 * self.ph)
        self.Jptpc = -self._advect((self.p[0] - self.p[1]),
                                   (self.del1 * self.u[0] + self.del2 * self.u[1]),
                                   (self.del1 * self.v[0] + self.del2 * self.v[1]))
        self.Jpxi = self._advect(self.xi, self.u, self.v)
        self.Jq = self._advect(self.q, self.u, self.v)
        self.Sph = np.einsum("ij,jkl->ikl", self.S, self.ph)

    def set_q1q2(self, q1, q2, check=False):
        self.q = np.vstack([q1[np.newaxis, :, :], q2[np.newaxis, :, :]])
        if check:
            np.testing.assert_allclose(self.q[0], q1)

    def set_U1U2(self, U1, U2):
        self.U1 = U1
        self.U2 = U2
        self.Ubg = np.array([U1, U2])


class LayeredModel(Model):
    """N-layer QG model (pyqg/layered_model.py:17-229)."""

    def __init__(self, beta=1.5e-11, nz=3, rd=15000.0, f=0.0001236812857687059, H=None, U=None,
                 V=None, rho=None, delta=None, **kwargs):
        self.beta = beta
        self.rd = rd
        self.delta = delta
        if U is None:
            U = (np.arange(nz) * 0.025)[::-1]
        if V is None:
            V = np.zeros(nz)
        if H is None:
            H = [500] + [1750 for _ in range(nz - 1)]
        if rho is None:
            rho = np.arange(nz) * 0.3 + 1025
        self.Ubg = np.array(U)
        self.Vbg = np.array(V)
        self.Hi = np.array(H)
        self.rhoi = np.array(rho)
        super().__init__(nz=nz, f=f, **kwargs)
        self.vertical_modes()
        self.set_q(1e-7 * np.vstack([np.random.randn(self.nx, self.ny)[np.newaxis]
                                     for _ in range(nz)]))

    def _initialize_stretching_matrix(self):
        """Stretching matrix S (layered_model.py:136-168)."""
        if self.nz == 2 and self.rd and self.delta:
            self.del1 = self.delta / (self.delta + 1.)
            self.del2 = (self.delta + 1.) ** -1
            self.Us = self.Ubg[0] - self.Ubg[1]
            self.F1 = self.rd ** -2 / (1. + self.delta)
            self.F2 = self.delta * self.F1
        self.S, _ = C.layered_stretching(self.nz, self.Hi, self.rhoi, self.g,
                                         getattr(self, "f2", None), self.rd, self.delta)

    def _initialize_background(self):
        """layered_model.py:170-209"""
        self.H = self.Hi.sum()
        if self.nz != 2:
            self.gpi = self.g * (self.rhoi[1:] - self.rhoi[:-1]) / self.rhoi[:-1]
            self.f2gpi = (self.f2 / self.gpi)[:, np.newaxis, np.newaxis]
            assert self.gpi.size == self.nz - 1, "Invalid size of gpi"
            assert np.all(self.gpi > 0.), "Buoyancy jump has negative sign!"
            assert self.Hi.size == self.nz, 'size of Hi does not match number of vertical levels nz'
            assert self.rhoi.size == self.nz, 'size of rhoi does not match number of vertical levels nz'
            assert self.Ubg.size == self.nz, 'size of Ubg does not match number of vertical levels nz'
            assert self.Vbg.size == self.nz, 'size of Vbg does not match number of vertical levels nz'
        else:
            self.f2gpi = np.array(self.rd ** -2 * (self.Hi[0] * self.Hi[1]) / self.H
                                  )[np.newaxis, np.newaxis]
        self._initialize_stretching_matrix()
        self.Qy = self.beta - np.dot(self.S, self.Ubg)
        self.Qx = np.dot(self.S, self.Vbg)
        self.ikQy = self.Qy[:, np.newaxis, np.newaxis] * 1j * self.k
        self.ilQx = self.Qx[:, np.newaxis, np.newaxis] * 1j * self.l

    def _initialize_inversion_matrix(self):
        """a = (S - kappa^2 I)^-1 per wavenumber, zero at kappa = 0
        (layered_model.py:211-229).  With three or more layers the fused steps
        solve (S - kappa^2 I) ph = qh with the tridiagonal S instead of reading
        the nz^2 table; the per-method phases keep using a."""
        self.a = C.layered_inversion(self.S, self.wv2)
        if self.nz >= 3:
            try:
                self._bucket.set_inversion_tridiagonal(self.S)
            except (_lib.QGKError, ValueError):
                pass  # an S the solve does not take: the fused steps read a

    def modal_projection(self, p, forward=True):
        """Modal amplitudes of p in the basis pmodes, or the inverse
        (model.py:307-317), as a solve at every wavenumber."""
        if forward:
            return np.linalg.solve(self.pmodes, p.reshape(self.nz, -1)).reshape(p.shape)
        return np.einsum("ik,k...->i...", self.pmodes, p)

    def _initialize_model_diagnostics(self):
        """qg_diagnostics.py:6-70 and layered_model.py:256-301"""
        self._add_device_diagnostics(Dg.QGDIAG)
        self._add_device_diagnostics(Dg.LAYERED_ONLY)

    def _calc_derived_fields(self):
        """Host fields of pyqg's diagnostic functions (layered_model.py:242-254)."""
        self.p = self.ifft(self.ph)
        self.xi = self.ifft(-self.wv2 * self.ph)
        self.Jpxi = self._advect(self.xi, self.u, self.v)
        self.Jq = self._advect(self.q, self.u, self.v)
        self.Sph = np.einsum("ij,jkl->ikl", self.S, self.ph)
        self.Sp = self.ifft(self.Sph)
        self.JSp = self._advect(self.Sp, self.u, self.v)
        self.phn = self.modal_projection(self.ph)


class BTModel(Model):
    """Single-layer (equivalent) barotropic model (pyqg/bt_model.py:6-104)."""

    def __init__(self, beta=0., rd=None, H=1., U=0., **kwargs):
        self.beta = beta
        self.rd = rd
        self.H = H
        self.Hi = np.array(H)[np.newaxis, ...]
        self.U = U
        self.nz = 1
        self.kd2 = rd ** -2 if rd else 0.
        super().__init__(**kwargs)
        self.set_q(1e-3 * np.random.rand(1, self.ny, self.nx))

    def _initialize_background(self):
        self.Qy = np.asarray(self.beta)[np.newaxis, ...]
        self.set_U(self.U)
        self.ikQy = self.Qy * 1j * self.k
        self.ilQx = 0.

    def _initialize_inversion_matrix(self):
        if self.rd:
            self.a = -(1. / (self.wv2 + self.kd2))[np.newaxis, np.newaxis, :, :]
        else:
            self.a = -self.wv2i[np.newaxis, np.newaxis, :, :]

    def set_U(self, U):
        self.Ubg = np.asarray(U)[np.newaxis, ...]

    def _diagnostics_due(self, t, tc):
        """No tavestart test in this model (bt_model.py:106-109, sqg_model.py:75-78)."""
        return (t >= self.dt) and (tc % self.taveints == 0)

    def _calc_cfl(self):
        return np.abs(np.hstack([self.u + self.Ubg, self.v])).max() * self.dt / self.dx

    def _calc_ke(self):
        return (.5 * self.spec_var(self.wv * self.ph)).sum()


class SQGModel(Model):
    """Surface QG model (pyqg/sqg_model.py:6-79)."""

    def __init__(self, beta=0., Nb=1., f_0=1., H=1., U=0., **kwargs):
        self.beta = beta
        self.Nb = Nb
        self.f_0 = f_0
        self.H = H
        self.Hi = np.array(H)[np.newaxis, ...]
        self.U = U
        self.nz = 1
        super().__init__(**kwargs)
        self.set_q(1e-3 * np.random.rand(1, self.ny, self.nx))

    def _initialize_background(self):
        self.Qy = np.asarray(self.beta)[np.newaxis, ...]
        self.set_U(self.U)
        self.ikQy = self.Qy * 1j * self.k
        self.ilQx = 0.

    def _initialize_inversion_matrix(self):
        self.a = np.asarray(self.f_0 / self.Nb * np.sqrt(self.wv2i))[np.newaxis, np.newaxis, :, :]

    def set_U(self, U):
        self.Ubg = np.asarray(U)[np.newaxis, ...]

    def _diagnostics_due(self, t, tc):
        """No tavestart test in this model (bt_model.py:106-109, sqg_model.py:75-78)."""
        return (t >= self.dt) and (tc % self.taveints == 0)

    def _calc_cfl(self):
        return np.abs(np.hstack([self.u + self.Ubg, self.v])).max() * self.dt / self.dx

    def _calc_ke(self):
        return (.5 * self.spec_var(self.wv * self.ph)).sum()
