"""Subgrid parameterizations (pyqg/parameterizations.py:1-360).

The class tree, composition (`a + b`) and weighting (`0.5 * a`) are pyqg's.
The built-in closures Smagorinsky, BackscatterBiharmonic and ZannaBolton2020
are evaluated by libqgk on the GPU from the model's current u, v and ph: as a
model's parameterization they write the forcing straight into the kernel's
du/dv or dq fields (sums and weights accumulate there), with no host round
trip; called directly, `p(m)` returns host arrays as pyqg's do.  RingForcing
draws host random numbers and runs on the host as in pyqg.
"""
from abc import ABC, abstractmethod
import ctypes

import numpy as np

from . import _lib


def _device_target(m):
    """(bucket, dx[M], Hi[M][nz]) when m is a pyqg_amd model, else None."""
    b = getattr(m, "_bucket", None)
    if b is None or not hasattr(m, "dx"):
        return None
    dx = np.full(b.M, float(m.dx))
    Hi = np.broadcast_to(np.asarray(getattr(m, "Hi", np.ones(m.nz)), dtype=np.float64).ravel(),
                         (m.nz,))
    return b, dx, np.ascontiguousarray(np.tile(Hi, (b.M, 1)))


def _require_device(m, name):
    t = _device_target(m)
    if t is None:
        raise TypeError(f"{name} runs on the GPU and needs a pyqg_amd model")
    if hasattr(m, "_flush"):
        m._flush()
    return t


def _capture_steps(b, dx, Hi, qp, uvp, w):
    """Record two phase-by-phase steps of bucket b with the closures qp (into dq)
    and uvp (into du, dv) as a HIP graph (qgk_capture_begin/end): invert,
    advection, friction, closures, time step with the AB weights w.  Two steps
    are one period of the tendency history's buffer swap."""
    L = _lib.lib
    _lib.check(L.qgk_capture_begin(b.h))
    try:
        for _ in range(2):
            for fn in (L.qgk_invert, L.qgk_advect, L.qgk_friction):
                _lib.check(fn(b.h))
            if uvp is not None:
                uvp._device_bucket(b, dx, Hi, 1.0, False)
                _lib.check(L.qgk_apply_uv_param(b.h))
            if qp is not None:
                qp._device_bucket(b, dx, Hi, 1.0, False)
                _lib.check(L.qgk_apply_q_param(b.h))
            _lib.check(L.qgk_timestep(b.h, *w))
    except Exception:
        L.qgk_capture_end(b.h, 0)
        raise
    _lib.check(L.qgk_capture_end(b.h, 1))


class Parameterization(ABC):
    """A subgrid parameterization (parameterizations.py:4-81)."""

    @abstractmethod
    def __call__(self, m):
        pass

    @property
    @abstractmethod
    def parameterization_type(self):
        pass

    def __add__(self, other):
        return CompositeParameterization(self, other)

    def __mul__(self, constant):
        return WeightedParameterization(self, constant)

    __rmul__ = __mul__

    # device evaluation into a bucket's parameterization fields
    def _device_capable(self):
        return False

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        raise NotImplementedError

    def _device_key(self):
        """Every value a device evaluation bakes into a captured step graph, at
        full precision (floats compare exactly); a kept graph is replayed only
        while this key is unchanged."""
        raise NotImplementedError

    def _device(self, m, weight=1.0, accumulate=False):
        b, dx, Hi = _require_device(m, type(self).__name__)
        self._device_bucket(b, dx, Hi, weight, accumulate)


class CompositeParameterization(Parameterization):
    """Sum of parameterizations of one type (parameterizations.py:83-100)."""

    def __init__(self, *params):
        assert len(set(p.parameterization_type for p in params)) == 1, \
            "all parameterizations must target the same variable (uv or q)"
        self.params = params

    @property
    def parameterization_type(self):
        return self.params[0].parameterization_type

    def __call__(self, m):
        return np.sum([f(m) for f in self.params], axis=0)

    def _device_capable(self):
        return all(p._device_capable() for p in self.params)

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        for i, p in enumerate(self.params):
            p._device_bucket(b, dx, Hi, weight, accumulate or i > 0)

    def _device_key(self):
        return ("Composite",) + tuple(p._device_key() for p in self.params)

    def __repr__(self):
        return f"CompositeParameterization{self.params}"


class WeightedParameterization(Parameterization):
    """A parameterization times a constant (parameterizations.py:102-117)."""

    def __init__(self, param, weight):
        self.param = param
        self.weight = weight

    @property
    def parameterization_type(self):
        return self.param.parameterization_type

    def __call__(self, m):
        return np.array(self.param(m)) * self.weight

    def _device_capable(self):
        return self.param._device_capable()

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        self.param._device_bucket(b, dx, Hi, weight * self.weight, accumulate)

    def _device_key(self):
        return ("Weighted", float(self.weight), self.param._device_key())

    def __repr__(self):
        return f"{self.weight} * {self.param}"


class UVParameterization(Parameterization):
    """Parameterization of the velocity tendencies (parameterizations.py:119-123)."""
    parameterization_type = 'uv_parameterization'


class QParameterization(Parameterization):
    """Parameterization of the PV tendency (parameterizations.py:125-130)."""
    parameterization_type = 'q_parameterization'


def _host_pair(b):
    shape = (b.M, b.nz, b.ny, b.nx)
    return np.empty(shape), np.empty(shape)


class Smagorinsky(UVParameterization):
    """Eddy viscosity nu = (Cs dx)^2 sqrt(2 (Sxx^2 + Syy^2 + 2 Sxy^2)) and the
    stress divergence du = 2 (d_x(nu Sxx) + d_y(nu Sxy)),
    dv = 2 (d_x(nu Sxy) + d_y(nu Syy)) (parameterizations.py:132-192), all
    derivatives spectral."""

    def __init__(self, constant=0.1):
        self.constant = constant

    def __call__(self, m, just_viscosity=False):
        b, dx, _ = _require_device(m, "Smagorinsky")
        if just_viscosity:
            nu = np.empty((b.M, b.nz, b.ny, b.nx))
            _lib.check(_lib.lib.qgk_smagorinsky_viscosity(b.h, _lib.ptr(dx),
                                                          float(self.constant), _lib.ptr(nu)))
            return nu[0] if b.M == 1 else nu
        du, dv = _host_pair(b)
        _lib.check(_lib.lib.qgk_param_smagorinsky(b.h, _lib.ptr(dx), float(self.constant), 1.0,
                                                  0, _lib.ptr(du), _lib.ptr(dv)))
        return (du[0], dv[0]) if b.M == 1 else (du, dv)

    def _device_capable(self):
        return True

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        _lib.check(_lib.lib.qgk_param_smagorinsky(b.h, _lib.ptr(dx), float(self.constant),
                                                  float(weight), int(bool(accumulate)), None,
                                                  None))

    def _device_key(self):
        return ("Smagorinsky", float(self.constant))

    def __repr__(self):
        return f"Smagorinsky(Cs={self.constant})"


class BackscatterBiharmonic(QParameterization):
    """Biharmonic Smagorinsky dissipation plus an energy-consistent backscatter
    of a fraction Cb of it (Jansen & Held 2014; parameterizations.py:194-240):
    dq = -lap(F^-1(lap^2 psi dx^2 nu)) - Cb lap^2 psi <psi diss>_H / (<psi lap^2 psi>_H + eps)."""

    def __init__(self, smag_constant=0.08, back_constant=0.99, eps=1e-32):
        self.smagorinsky = Smagorinsky(smag_constant)
        self.back_constant = back_constant
        self.eps = eps

    def __call__(self, m):
        b, dx, Hi = _require_device(m, "BackscatterBiharmonic")
        dq = np.empty((b.M, b.nz, b.ny, b.nx))
        self._run(b, dx, Hi, 1.0, 0, _lib.ptr(dq))
        return dq[0] if b.M == 1 else dq

    def _run(self, b, dx, Hi, weight, accumulate, out):
        _lib.check(_lib.lib.qgk_param_backscatter(
            b.h, _lib.ptr(dx), _lib.ptr(Hi), float(self.smagorinsky.constant),
            float(self.back_constant), float(self.eps), float(weight), int(bool(accumulate)),
            out))

    def _device_capable(self):
        return True

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        self._run(b, dx, Hi, weight, accumulate, None)

    def _device_key(self):
        return ("BackscatterBiharmonic", float(self.smagorinsky.constant),
                float(self.back_constant), float(self.eps))

    def __repr__(self):
        return (f"BackscatterBiharmonic(Cs={self.smagorinsky.constant}, "
                f"Cb={self.back_constant})")


class ZannaBolton2020(UVParameterization):
    """Equation-discovered velocity closure of Zanna & Bolton 2020, Eq. 6
    (parameterizations.py:242-289): from the relative vorticity, shearing and
    stretching deformations, du = k F^-1(ik (S/2 - rv*sh) + il rv*st),
    dv = k F^-1(il (S/2 + rv*sh) + ik rv*st), S = F(rv^2 + sh^2 + st^2)."""

    def __init__(self, constant=-46761284):
        self.constant = constant

    def __call__(self, m):
        b, _, _ = _require_device(m, "ZannaBolton2020")
        du, dv = _host_pair(b)
        _lib.check(_lib.lib.qgk_param_zb2020(b.h, float(self.constant), 1.0, 0, _lib.ptr(du),
                                             _lib.ptr(dv)))
        return (du[0], dv[0]) if b.M == 1 else (du, dv)

    def _device_capable(self):
        return True

    def _device_bucket(self, b, dx, Hi, weight, accumulate):
        _lib.check(_lib.lib.qgk_param_zb2020(b.h, float(self.constant), float(weight),
                                             int(bool(accumulate)), None, None))

    def _device_key(self):
        return ("ZannaBolton2020", float(self.constant))

    def __repr__(self):
        return f"ZannaBolton2020(κ={self.constant:.2e})"


class RingForcing(QParameterization):
    """Stochastic PV forcing on a ring of wavenumbers k_in < |k| <= k_out
    (Uchida et al. 2023; parameterizations.py:291-360): white noise in the
    ring, transformed to real space, mean removed, scaled to unit mean
    absolute value and by mag_noise_forc; 'all', 'surf' or 'bottom' layers.
    Host random numbers, as in pyqg."""

    def __init__(self, k_in_forc=0, k_out_forc=0, mag_noise_forc=0, layers='all'):
        self.k_in_forc = k_in_forc
        self.k_out_forc = k_out_forc
        self.mag_noise_forc = mag_noise_forc
        self.layers = layers

    def __call__(self, m):
        nhx, nhy = m.wv.shape
        wvx = np.sqrt(m.k ** 2. + m.l ** 2.)
        mask = np.ones_like(wvx)
        mask[wvx <= self.k_in_forc] = 0.
        mask[wvx > self.k_out_forc] = 0.
        ring_hat = mask * (np.random.randn(nhx, nhy) + 1j * np.random.randn(nhx, nhy))
        full = m.ifft(np.broadcast_to(ring_hat[np.newaxis], (m.nz, nhx, nhy)).copy())
        if self.layers == 'all':
            ring = full
        else:
            ring = np.zeros_like(m.u)
            if self.layers == 'surf':
                ring[0] = full[0]
            elif self.layers == 'bottom':
                ring[-1] = full[-1]
        ring = ring - ring.mean(axis=(-1, -2))[:, np.newaxis, np.newaxis]
        ring = ring / np.abs(ring).mean(axis=(-1, -2))[:, np.newaxis, np.newaxis]
        return self.mag_noise_forc * ring

    def __repr__(self):
        return (f"RingForcing(k_in={self.k_in_forc}, k_out={self.k_out_forc}, "
                f"mag={self.mag_noise_forc}, layers={self.layers})")
