"""Ensembles: many independent pyqg models stepped together on one GPU.

pyqg has no ensemble concept (users build separate Model objects); here the
members of one grid size share one libqgk handle ("bucket"), so one fused step
of every member is two kernel launches per bucket.  Members of different nx
go to different buckets; all buckets are stepped in one call.

Each member follows exactly the per-model semantics of the reference
(pyqg/model.py:406-436 with kernel.pyx:312-526): the same constants as the
corresponding QGModel / LayeredModel would build, the same AB start-up and the
same accumulated time t.
"""
import json
import os

import numpy as np

from . import _lib
from . import constants as C
from ._lib import Bucket
from .diagnostics import param_block
from .errors import DiagnosticNotFilledError


def _as_array(v, n):
    a = np.asarray(v, dtype=np.float64)
    return np.full(n, float(a)) if a.ndim == 0 else a.astype(np.float64)


class Ensemble:
    """A set of independent models with a common time step dt."""

    def __init__(self, dt):
        self._dt = float(dt)
        self.t = 0.0
        self.tc = 0
        self.ablevel = 0
        self.buckets = []      # Bucket handles, one per (nz, nx)
        self.members = []      # member index -> (bucket index, position in bucket)
        self.member_nx = []
        self.diag_params = []  # per bucket: (M, plen) constants of the device diagnostics
        self.diag_count = {}
        self.closure_dx = []   # per bucket: grid spacing [M] and layer thicknesses [M][nz]
        self.closure_Hi = []
        self.failed = set()    # members masked after a failed health check
        self.cfl = None        # per member: CFL number at the last health check

    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, dt):
        self._dt = float(dt)
        self.ablevel = 0

    @property
    def nmembers(self):
        return len(self.members)

    # ------------------------------------------------------------ builders
    @classmethod
    def qg(cls, nx, beta=1.5e-11, rd=15000.0, delta=0.25, rek=5.787e-7, L=1e6, H1=500.,
           U1=0.025, U2=0.0, dt=3600., filterfac=23.6, nmembers=None, full_state=False,
           q_param=False, uv_param=False):
        """Two-layer QGModels (qg_model.py:60-161).  Every argument may be a
        scalar or one value per member; members are bucketed by nx.  full_state
        keeps every intermediate field (needed by the per-phase operations and
        the closure pipeline); q_param allocates dq / dqh, uv_param du, dv,
        duh, dvh."""
        if nmembers is None:
            sizes = [np.size(v) for v in (nx, beta, rd, delta, rek, L, U1, U2)]
            nmembers = max(sizes)
        n = int(nmembers)
        nxs = np.full(n, int(nx)) if np.ndim(nx) == 0 else np.asarray(nx, dtype=int)
        p = {k: _as_array(v, n) for k, v in dict(beta=beta, rd=rd, delta=delta, rek=rek, L=L,
                                                  H1=H1, U1=U1, U2=U2,
                                                  filterfac=filterfac).items()}
        ens = cls(dt)
        ens.member_nx = list(nxs)
        ens.members = [None] * n
        for size in sorted(set(nxs.tolist())):
            idx = np.nonzero(nxs == size)[0]
            flags = ((_lib.FULL_STATE if full_state else 0) | (_lib.HAS_Q_PARAM if q_param else 0)
                     | (_lib.HAS_UV_PARAM if uv_param else 0))
            b = Bucket(len(idx), 2, size, size, flags)
            nk = size // 2 + 1
            kk = np.empty((len(idx), nk))
            ll = np.empty((len(idx), size))
            k2l2 = np.empty((len(idx), size, nk))
            filt = np.empty((len(idx), size, nk))
            Ubg = np.empty((len(idx), 2))
            Qy = np.empty((len(idx), 2))
            dparams = []
            for j, m in enumerate(idx):
                g = C.grid(size, size, p["L"][m], p["L"][m])
                kk[j], ll[j], k2l2[j] = g["kk"], g["ll"], g["k2l2"]
                filt[j] = C.spectral_filter(g["k"], g["l"], g["dx"], g["dy"], p["filterfac"][m])
                F1, F2, Qy[j], Ubg[j] = C.qg_background(p["beta"][m], p["rd"][m], p["delta"][m],
                                                        p["U1"][m], p["U2"][m])
                b.set_inversion_2layer(F1, F2, member=j)
                b.set_rek(p["rek"][m], member=j)
                d = p["delta"][m]
                Hi = np.array([p["H1"][m], p["H1"][m] / d])
                dparams.append(param_block(2, Hi, Hi.sum(), S=[[-F1, F1], [F2, -F2]],
                                           del1=d / (d + 1.), del2=(d + 1.) ** -1,
                                           rd=p["rd"][m], U=p["U1"][m] - p["U2"][m]))
                ens.members[m] = (len(ens.buckets), j)
            ens._set_table(b, "kk", kk)
            ens._set_table(b, "ll", ll)
            ens._set_table(b, "k2l2", k2l2)
            ens._set_table(b, "filtr", filt)
            ens._set_table(b, "Ubg", Ubg)
            ens._set_table(b, "Qy", Qy)
            ens.buckets.append(b)
            ens.diag_params.append(np.array(dparams))
            ens.closure_dx.append(np.array([p["L"][m] / size for m in idx]))
            ens.closure_Hi.append(np.array([[p["H1"][m], p["H1"][m] / p["delta"][m]]
                                            for m in idx]))
        ens.params = p
        return ens

    @classmethod
    def layered(cls, nmembers, nx, nz, H, rho, U, L=1e6, beta=1.5e-11, rek=5.787e-7,
                f=0.0001236812857687059, g=9.81, dt=3600., filterfac=23.6, full_state=False):
        """nmembers LayeredModels with one shared stratification
        (layered_model.py:72-229); the inversion table is shared by all."""
        n, nz = int(nmembers), int(nz)
        Hi, rhoi, Ubg = (np.asarray(v, dtype=np.float64) for v in (H, rho, U))
        gr = C.grid(nx, nx, L, L)
        S, _ = C.layered_stretching(nz, Hi, rhoi, g, f * f)
        Qy = beta - np.dot(S, Ubg)
        a = C.layered_inversion(S, gr["wv2"])
        ens = cls(dt)
        b = Bucket(n, nz, nx, nx, _lib.FULL_STATE if full_state else 0)
        b.set_const("kk", gr["kk"])
        b.set_const("ll", gr["ll"])
        b.set_const("k2l2", gr["k2l2"])
        b.set_const("filtr", C.spectral_filter(gr["k"], gr["l"], gr["dx"], gr["dy"], filterfac))
        b.set_const("a", a)
        if nz >= 3:
            try:
                b.set_inversion_tridiagonal(S)  # fused steps solve instead of reading a
            except (_lib.QGKError, ValueError):
                pass  # an S the solve does not take: the fused steps read a
        b.set_const("Ubg", Ubg)
        b.set_const("Qy", Qy)
        b.set_rek(rek)
        ens.buckets = [b]
        ens.members = [(0, j) for j in range(n)]
        ens.member_nx = [nx] * n
        ens.S, ens.Qy, ens.a = S, Qy, a
        gpi = g * (rhoi[1:] - rhoi[:-1]) / rhoi[:-1]
        kdi2, pmodes = C.vertical_modes(S, Hi)
        blk = param_block(nz, Hi, Hi.sum(), S=S, f2gpi=f * f / gpi, pmodes=pmodes, kdi2=kdi2)
        ens.diag_params = [np.tile(blk, (n, 1))]
        ens.closure_dx = [np.full(n, L / nx)]
        ens.closure_Hi = [np.tile(Hi, (n, 1))]
        return ens

    @staticmethod
    def _set_table(b, name, per_member):
        """Upload one shared table when every member's is identical, else one
        table per member."""
        if all(np.array_equal(per_member[0], x) for x in per_member[1:]):
            b.set_const(name, per_member[0])
        else:
            b.set_const_all(name, per_member)

    # ------------------------------------------------------------ state
    def _bucket_members(self, bi):
        return [m for m, (b, _) in enumerate(self.members) if b == bi]

    def set_q(self, q):
        """Initial PV of every member: a sequence of (nz, ny, nx) arrays (or one
        array with a leading member axis when all members share a size).  Sets
        qh = rfft2(q) as the reference's q setter does (kernel.pyx:576-582)."""
        for bi, b in enumerate(self.buckets):
            mem = self._bucket_members(bi)
            arr = np.stack([np.asarray(q[m], dtype=np.float64) for m in mem])
            b.put("q", arr)
            _lib.check(_lib.lib.qgk_q_to_qh(b.h))

    def get(self, field, member=None):
        """A field ("q", "qh", "dqhdt_p", ...) of one member, or a list for all."""
        out = [b.get(field) for b in self.buckets]
        if member is not None:
            bi, j = self.members[member]
            return out[bi][j]
        return [out[bi][j] for (bi, j) in self.members]

    # ------------------------------------------------------------ stepping
    def step(self, nsteps=1, use_graph=False, twrite=None, on_failure="mask"):
        """nsteps fused steps of every member.  With twrite, every member's
        health is checked whenever tc reaches a multiple of twrite, as the
        reference asserts CFL < 1 at its status steps (model.py:564-574);
        check_members says what happens to a member that fails."""
        nsteps = int(nsteps)
        while nsteps > 0:
            n = nsteps
            if twrite:
                n = min(n, int(twrite) - self.tc % int(twrite))
            self._step(n, use_graph)
            nsteps -= n
            if twrite and self.tc % int(twrite) == 0:
                self.check_members(on_failure)

    def _step(self, nsteps, use_graph):
        if nsteps <= 0:
            return
        coeffs = np.empty((nsteps, 3))
        for s in range(nsteps):
            w, self.ablevel = C.ab_weights(self.ablevel, self._dt)
            coeffs[s] = w
        _lib.step_buckets(self.buckets, coeffs, use_graph)
        for _ in range(nsteps):
            self.t += self._dt
        self.tc += nsteps

    def member_cfl(self):
        """CFL number max(|u + Ubg|, |v|) dt / dx of every member's current
        state (qg_diagnostics.py:76-79), computed on the device for all members
        at once; NaN or inf for a member whose state is not finite."""
        cfl = np.empty(self.nmembers)
        for bi, (b, dx) in enumerate(zip(self.buckets, self.closure_dx)):
            sp = b.member_speed()
            for j, m in enumerate(self._bucket_members(bi)):
                cfl[m] = sp[j] * self._dt / dx[j]
        return cfl

    def check_members(self, on_failure="mask"):
        """Health check of every member: a member fails when its CFL number is
        not below 1 or its state is not finite.  on_failure="mask" zeroes a
        failed member's state (it rests at q = 0 and no longer feeds NaNs into
        member sums) and records it in self.failed; "raise" raises the
        reference's AssertionError instead.  Returns the members that failed
        in this check."""
        if on_failure not in ("mask", "raise"):
            raise ValueError("on_failure must be 'mask' or 'raise'")
        self.cfl = cfl = self.member_cfl()
        bad = [m for m in range(self.nmembers)
               if m not in self.failed and not (np.isfinite(cfl[m]) and cfl[m] < 1.0)]
        if bad and on_failure == "raise":
            raise AssertionError(f"CFL condition violated in members {bad} "
                                 f"(CFL {[float(cfl[m]) for m in bad]})")
        for m in bad:
            bi, j = self.members[m]
            self.buckets[bi].mask_member(j)
            self.failed.add(m)
        return bad

    def autotune(self, reps=5, cache_path=None):
        """Pick the fastest register radix of the fused row and column kernels
        for every bucket (pyqg_amd._lib.Bucket.autotune); results are cached
        per device and bucket shape in a JSON file (default
        ~/.cache/pyqg_amd/autotune.json; pass cache_path=False to disable)."""
        if cache_path is None:
            cache_path = os.path.join(os.path.expanduser("~"), ".cache", "pyqg_amd",
                                      "autotune.json")
        cache = {}
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    cache = json.load(f)
            except (OSError, ValueError):
                cache = {}
        dev = _lib.device_name()
        report = {}
        for b in self.buckets:
            key = f"{dev}|nz={b.nz}|ny={b.ny}|nx={b.nx}|M={b.M}"
            if key in cache:
                b.set_variant(*cache[key]["variant"])
                report[key] = cache[key]
                continue
            variant, ms_row, ms_col = b.autotune(reps)
            cache[key] = report[key] = {"variant": list(variant), "ms_row": ms_row,
                                        "ms_col": ms_col}
        if cache_path:
            try:
                os.makedirs(os.path.dirname(cache_path), exist_ok=True)
                with open(cache_path, "w") as f:
                    json.dump(cache, f, indent=1)
            except OSError:
                pass
        return report

    def step_parameterized(self, nsteps, parameterization):
        """nsteps steps with a built-in closure (pyqg_amd.parameterizations) in
        every member, in the reference's phase order (model.py:406-436):
        inversion, advection, friction, the closure evaluated and added on the
        device, time step.  The ensemble needs full_state and q_param or
        uv_param matching the closure."""
        p = parameterization
        if not p._device_capable():
            raise ValueError(f"{p!r} has no device evaluation")
        nsteps = int(nsteps)
        while nsteps > 0 and self.ablevel < 2:
            self._param_step(p)       # AB start-up
            nsteps -= 1
        if nsteps >= 2:
            # once the AB weights are constant, every bucket replays a HIP graph
            # of two of these steps (the same kernels, bit for bit)
            from .parameterizations import _capture_steps
            uv = p.parameterization_type == 'uv_parameterization'
            w, self.ablevel = C.ab_weights(self.ablevel, self._dt)
            key = (p._device_key(), tuple(w),
                   tuple(np.asarray(x, dtype=float).tobytes() for x in self.closure_dx),
                   tuple(np.asarray(x, dtype=float).tobytes() for x in self.closure_Hi))
            L = _lib.lib

            def replayable():
                return all(L.qgk_graph_replay(b.h, 0) == 0 for b in self.buckets)

            ok = getattr(self, "_graph_key", None) == key
            if not ok or not replayable():
                # allocations and uploads before a capture; the history phase
                # of a kept graph
                self._param_step(p)
                nsteps -= 1
                ok = ok and replayable()
            if not ok and nsteps >= 2:
                self._graph_key = None
                for b, dx, Hi in zip(self.buckets, self.closure_dx, self.closure_Hi):
                    _capture_steps(b, np.ascontiguousarray(dx), np.ascontiguousarray(Hi),
                                   None if uv else p, p if uv else None, w)
                self._graph_key = key
                ok = True
            if ok and nsteps >= 2:
                reps = nsteps // 2
                for b in self.buckets:
                    _lib.check(L.qgk_graph_replay(b.h, reps))
                for _ in range(2 * reps):
                    self.t += self._dt
                self.tc += 2 * reps
                nsteps -= 2 * reps
        for _ in range(nsteps):
            self._param_step(p)

    def _param_step(self, p):
        """One step of every bucket, phase by phase, with the device closure p."""
        uv = p.parameterization_type == 'uv_parameterization'
        L = _lib.lib
        w, self.ablevel = C.ab_weights(self.ablevel, self._dt)
        for b, dx, Hi in zip(self.buckets, self.closure_dx, self.closure_Hi):
            for fn in (L.qgk_invert, L.qgk_advect, L.qgk_friction):
                _lib.check(fn(b.h))
            p._device_bucket(b, np.ascontiguousarray(dx), np.ascontiguousarray(Hi), 1.0, False)
            _lib.check((L.qgk_apply_uv_param if uv else L.qgk_apply_q_param)(b.h))
            _lib.check(L.qgk_timestep(b.h, *w))
        self.t += self._dt
        self.tc += 1

    # ------------------------------------------------------------ diagnostics
    def accumulate_diagnostics(self, names=("KEspec",)):
        """Add the named diagnostics (pyqg's registry names) of the current
        state of every member, on the device.  As in the reference they are
        evaluated inside the step, after the inversion, advection and friction
        of the current state and before the time step (model.py:406-436); the
        ensemble's state is unchanged.  Needs full_state=True."""
        names = list(names)
        w, _ = C.ab_weights(self.ablevel, self._dt)
        for b, prm in zip(self.buckets, self.diag_params):
            if not b.flags & _lib.FULL_STATE:
                raise ValueError("diagnostics need an ensemble built with full_state=True")
            if not getattr(b, "_diag_params_set", False):
                for j in range(b.M):
                    b.diag_set_params(j, prm[j])
                b._diag_params_set = True
            for fn in (_lib.lib.qgk_invert, _lib.lib.qgk_advect, _lib.lib.qgk_friction):
                _lib.check(fn(b.h))
            b.diag_accumulate(names, self._dt, w)
        for n in names:
            self.diag_count[n] = self.diag_count.get(n, 0) + 1

    def get_diagnostic(self, name, member=None, mean=False):
        """Time mean of one diagnostic: of one member, of every member (a list),
        or (mean=True) the mean over members of equal grid size (one array per
        bucket, in bucket order)."""
        cnt = self.diag_count.get(name, 0)
        if cnt == 0:
            raise DiagnosticNotFilledError(name)
        if member is not None:
            bi, j = self.members[member]
            return self.buckets[bi].diag_read(name, member=j) / cnt
        if mean:
            return [b.diag_read(name, member=-2) / (cnt * b.M) for b in self.buckets]
        per = [b.diag_read(name, member=-1) / cnt for b in self.buckets]
        return [per[bi][j] for (bi, j) in self.members]

    def mean_model(self, bucket=0):
        """The time-mean diagnostics of one bucket, averaged over its members,
        as an object pyqg's comparison tools accept in place of a model
        (diagnostic_tools.diagnostic_differences / diagnostic_similarities,
        pyqg/diagnostic_tools.py:128-262): a registry of the diagnostics
        accumulated so far with their dims, sums_with and skip_comparison
        entries, get_diagnostic, nz and the grid attributes calc_ispec reads."""
        return _MeanModel(self, bucket)

    def reset_diagnostics(self):
        for b in self.buckets:
            if b.flags & _lib.FULL_STATE and self.diag_count:
                b.diag_reset()
        self.diag_count = {}

    def model_days(self, nsteps):
        """Simulated model-days of nsteps steps summed over members."""
        return self.nmembers * nsteps * self._dt / 86400.

    def close(self):
        for b in self.buckets:
            b.close()
        self.buckets = []


class _MeanModel:
    """Member-mean time averages of one ensemble bucket, shaped like a model
    for the diagnostic comparison tools (see Ensemble.mean_model)."""

    def __init__(self, ens, bucket):
        from . import diagnostic_tools as tools
        from .diagnostics import CORE, LAYERED_ONLY, QG_ONLY, QGDIAG
        b = ens.buckets[bucket]
        L = float(ens.closure_dx[bucket][0]) * b.nx
        g = tools.grid(b.nx, L)
        for a in ("kk", "ll", "dk", "dl", "wv", "wv2", "M", "nx", "ny"):
            setattr(self, a, getattr(g, a))
        self.nz = b.nz
        self._ens, self._bucket = ens, bucket
        table = dict(CORE, **QGDIAG, **(QG_ONLY if b.nz == 2 else LAYERED_ONLY))
        self.diagnostics = {}
        for name, (desc, units, dims, extra) in table.items():
            if ens.diag_count.get(name, 0) == 0:
                continue
            entry = dict(description=desc, units=units, dims=dims, **extra)
            entry["function"] = lambda m, n=name: m.get_diagnostic(n)
            self.diagnostics[name] = entry

    def get_diagnostic(self, name):
        return self._ens.get_diagnostic(name, mean=True)[self._bucket]

