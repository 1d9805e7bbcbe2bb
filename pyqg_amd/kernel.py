"""Drop-in replacement of pyqg's `PseudoSpectralKernel` backed by HIP kernels.

Same constructor, methods, properties and error behaviour as
`cdef class PseudoSpectralKernel` (pyqg/kernel.pyx:32-637); the state lives in
HBM inside a libqgk handle and every method is a call through the C-ABI
(include/pyqg_amd.h).  pyqg's `Model` (pyqg/model.py:22) can inherit from this
class unchanged.

Host mirrors.  A property getter copies the field from the device into a host
array and returns that array; it stays "exposed" until the next kernel method,
which first uploads every exposed array, so writes made through a getter
(pyqg/tests/test_model.py:125-128, 174, 322) reach the kernel as they do in the
reference.  A reference kept across a kernel method no longer aliases the state
(as with the reference's numpy.fft build, which rebinds its arrays).
"""
import numpy as np

from . import _lib
from ._lib import Bucket
from .constants import ab_weights

DTYPE_real = np.float64
DTYPE_com = np.complex128

# fields each method writes on the device
_WRITES = {
    "invert": ("ph", "uh", "vh", "u", "v"),
    "advect": ("uq", "vq", "uqh", "vqh", "dqhdt"),
    "friction": ("dqhdt",),
    "q_param": ("dq", "dqh", "dqhdt"),
    "uv_param": ("du", "dv", "duh", "dvh", "dqhdt"),
    "timestep": ("qh", "q", "dqhdt_p", "dqhdt_pp"),
    "q_set": ("qh",),
    "qh_set": ("q",),
}


def _typed(arr, dtype, ndim, name):
    """The reference's typed-buffer arguments: ndarray of exact dtype and ndim."""
    if not isinstance(arr, np.ndarray):
        raise TypeError(f"Argument '{name}' has incorrect type (expected numpy.ndarray, "
                        f"got {type(arr).__name__})")
    if arr.dtype != dtype:
        raise ValueError(f"Buffer dtype mismatch, expected '{np.dtype(dtype).name}' but got "
                         f"'{arr.dtype.name}'")
    if arr.ndim != ndim:
        raise ValueError(f"Buffer has wrong number of dimensions (expected {ndim}, got "
                         f"{arr.ndim})")
    return arr


class PseudoSpectralKernel:
    """Pseudo-spectral QG kernel on one MI355X (FP64)."""

    def __init__(self, nz, ny, nx, fftw_num_threads=1, has_q_param=0, has_uv_param=0):
        nz, ny, nx = int(nz), int(ny), int(nx)
        self.nz, self.ny, self.nx = nz, ny, nx
        self.nl = ny
        self.nk = int(nx / 2 + 1)
        flags = _lib.FULL_STATE
        if has_q_param:
            flags |= _lib.HAS_Q_PARAM
        if has_uv_param:
            flags |= _lib.HAS_UV_PARAM
        self._bucket = Bucket(1, nz, ny, nx, flags)
        self._has_q = bool(has_q_param)
        self._has_uv = bool(has_uv_param)
        self._host = {}       # field -> host mirror (nz, ...)
        self._exposed = set()  # fields whose host mirror may have been written
        self._valid = set()    # fields whose host mirror equals the device copy
        # constants (kernel.pyx:122-127, 185-192)
        pending = self.__dict__.pop("_pending_consts", {})
        self._kk_v = np.zeros(self.nk)
        self._ll_v = np.zeros(self.nl)
        self._ik_v = np.zeros(self.nk, DTYPE_com)
        self._il_v = np.zeros(self.nl, DTYPE_com)
        self._k2l2_v = np.zeros((self.nl, self.nk))
        self._a_v = np.zeros((nz, nz, self.nl, self.nk), DTYPE_com)
        self._Ubg_v = np.zeros(nz)
        self._Qy_v = np.zeros(nz)
        self._ikQy_v = np.zeros((nz, self.nk), DTYPE_com)
        self._filtr_v = np.ones((self.nl, self.nk))
        self._const_exposed = set()
        self._rek = 0.0
        self._dt = 0.0
        self._t = 0.0
        self._tc = 0
        self._ablevel = 0
        self.num_threads = fftw_num_threads
        for name, val in pending.items():
            setattr(self, name, val)

    # ------------------------------------------------------------------ state
    def _field_shape(self, f):
        return ((self.nz, self.ny, self.nx) if f in _lib.REAL_FIELDS
                else (self.nz, self.nl, self.nk))

    def _check_alloc(self, f):
        if f in ("dq", "dqh") and not self._has_q:
            raise AttributeError(f"'{type(self).__name__}' object has no attribute '{f}'")
        if f in ("du", "dv", "duh", "dvh") and not self._has_uv:
            raise AttributeError(f"'{type(self).__name__}' object has no attribute '{f}'")

    def _get_field(self, f):
        self._check_alloc(f)
        if f not in self._valid:
            if f not in self._host:
                dtype = DTYPE_real if f in _lib.REAL_FIELDS else DTYPE_com
                self._host[f] = np.empty(self._field_shape(f), dtype)
            self._bucket.get(f, self._host[f].reshape((1,) + self._field_shape(f)))
            self._valid.add(f)
        self._exposed.add(f)
        return self._host[f]

    def _flush(self):
        """Upload every host mirror handed out since the last kernel call."""
        for f in self._exposed:
            self._bucket.put(f, self._host[f])
        self._exposed.clear()
        for c in self._const_exposed:
            self._upload_const(c)
        self._const_exposed.clear()

    def _wrote(self, key):
        for f in _WRITES[key]:
            self._valid.discard(f)

    def _invalidate_all(self):
        self._valid.clear()

    # -------------------------------------------------------------- constants
    def _upload_const(self, c):
        b = self.__dict__.get("_bucket")
        if b is None:
            return
        if c == "kk":
            b.set_const("kk", self._kk_v)
        elif c == "ll":
            b.set_const("ll", self._ll_v)
        elif c == "k2l2":
            b.set_const("k2l2", self._k2l2_v)
        elif c == "a":
            b.set_const("a", np.ascontiguousarray(self._a_v.real))
        elif c == "Ubg":
            b.set_const("Ubg", self._Ubg_v)
        elif c == "Qy":
            b.set_const("Qy", self._Qy_v)
        elif c == "filtr":
            b.set_const("filtr", self._filtr_v)
        elif c == "rek":
            b.set_rek(self._rek)

    def _pre_init(self, name, value):
        """Setters called before __init__ (LayeredModel sets Ubg first,
        layered_model.py:120) are replayed once the handle exists."""
        if "_bucket" not in self.__dict__:
            self.__dict__.setdefault("_pending_consts", {})[name] = value
            return True
        return False

    # ------------------------------------------------------- generic transforms
    def _empty_real(self):
        return np.zeros((self.nz, self.ny, self.nx), dtype=DTYPE_real)

    def _empty_com(self):
        return np.zeros((self.nz, self.nl, self.nk), dtype=DTYPE_com)

    def fft(self, v):
        """Generic r2c FFT of a (nz, ny, nx) real array (kernel.pyx:291-299)."""
        v = _typed(v, DTYPE_real, 3, "v")
        if v.shape != (self.nz, self.ny, self.nx):
            raise ValueError(f"fft: expected shape {(self.nz, self.ny, self.nx)}, got {v.shape}")
        src = np.ascontiguousarray(v)
        out = np.empty((self.nz, self.nl, self.nk), DTYPE_com)
        _lib.check(_lib.lib.qgk_fft(self._bucket.h, self.nz, _lib.ptr(src), _lib.ptr(out)))
        return out

    def ifft(self, v):
        """Generic c2r FFT of a (nz, nl, nk) complex array (kernel.pyx:301-309)."""
        v = _typed(v, DTYPE_com, 3, "v")
        if v.shape != (self.nz, self.nl, self.nk):
            raise ValueError(f"ifft: expected shape {(self.nz, self.nl, self.nk)}, got {v.shape}")
        src = np.ascontiguousarray(v)
        out = np.empty((self.nz, self.ny, self.nx), DTYPE_real)
        _lib.check(_lib.lib.qgk_ifft(self._bucket.h, self.nz, _lib.ptr(src), _lib.ptr(out)))
        return out

    # ----------------------------------------------------------------- phases
    def _invert(self):
        """ph = a qh; uh, vh = -il ph, ik ph; u, v = irfft2 (kernel.pyx:312-355)."""
        self._flush()
        _lib.check(_lib.lib.qgk_invert(self._bucket.h))
        self._wrote("invert")

    def _do_advection(self):
        """uq, vq, their rfft2 and dqhdt = -(ik uqh + il vqh + ikQy ph) (kernel.pyx:357-395)."""
        self._flush()
        _lib.check(_lib.lib.qgk_advect(self._bucket.h))
        self._wrote("advect")

    def _do_friction(self):
        """Bottom drag on the lowest layer's tendency (kernel.pyx:444-461)."""
        self._flush()
        _lib.check(_lib.lib.qgk_friction(self._bucket.h))
        self._wrote("friction")

    def _device_parameterization(self, p):
        """A built-in closure (pyqg_amd.parameterizations) that writes its
        forcing into the device fields itself, or None."""
        dev = getattr(p, "_device_capable", None)
        return p if dev is not None and dev() and hasattr(self, "dx") else None

    def _do_q_subgrid_parameterization(self):
        """dqhdt += rfft2(q_parameterization(self)) (kernel.pyx:424-442).
        Built-in closures evaluate on the device into dq directly."""
        p = self._device_parameterization(self.q_parameterization)
        if p is not None:
            self._check_alloc("dq")
            self._flush()
            p._device(self)
            _lib.check(_lib.lib.qgk_apply_q_param(self._bucket.h))
            self._wrote("q_param")
            return
        dq = self.q_parameterization(self)
        self._check_alloc("dq")
        dq = np.ascontiguousarray(np.broadcast_to(np.asarray(dq, dtype=DTYPE_real),
                                                  (self.nz, self.ny, self.nx)))
        self._flush()
        _lib.check(_lib.lib.qgk_add_q_param(self._bucket.h, _lib.ptr(dq)))
        self._wrote("q_param")

    def _do_uv_subgrid_parameterization(self):
        """dqhdt += -il rfft2(du) + ik rfft2(dv) (kernel.pyx:397-422).
        Built-in closures evaluate on the device into du, dv directly."""
        p = self._device_parameterization(self.uv_parameterization)
        if p is not None:
            self._check_alloc("du")
            self._flush()
            p._device(self)
            _lib.check(_lib.lib.qgk_apply_uv_param(self._bucket.h))
            self._wrote("uv_param")
            return
        du, dv = self.uv_parameterization(self)
        self._check_alloc("du")
        shape = (self.nz, self.ny, self.nx)
        du = np.ascontiguousarray(np.broadcast_to(np.asarray(du, dtype=DTYPE_real), shape))
        dv = np.ascontiguousarray(np.broadcast_to(np.asarray(dv, dtype=DTYPE_real), shape))
        self._flush()
        _lib.check(_lib.lib.qgk_add_uv_param(self._bucket.h, _lib.ptr(du), _lib.ptr(dv)))
        self._wrote("uv_param")

    def _ab_weights(self):
        """AB1 / AB2 / AB3 weights and the ablevel update (kernel.pyx:482-500)."""
        w, self._ablevel = ab_weights(self._ablevel, self._dt)
        return w

    def _forward_timestep(self):
        """AB update with the exponential filter and history shift (kernel.pyx:463-526)."""
        self._flush()
        dt1, dt2, dt3 = self._ab_weights()
        _lib.check(_lib.lib.qgk_timestep(self._bucket.h, dt1, dt2, dt3))
        self._wrote("timestep")
        self._tc += 1
        self._t += self._dt

    def _fused_steps(self, nsteps):
        """nsteps unparameterized steps of Model._step_forward's kernel phases
        (invert, advection, friction, forward timestep) in the fused two-kernel
        form.  The last step runs phase by phase so ph, u, v, uq, vq, uqh, vqh
        and dqhdt are left exactly as the per-method path leaves them."""
        nsteps = int(nsteps)
        if nsteps <= 0:
            return
        self._flush()
        if nsteps > 1:
            coeffs = np.array([self._ab_weights() for _ in range(nsteps - 1)])
            self._bucket.step(coeffs)
            self._invalidate_all()
            self._tc += nsteps - 1
            for _ in range(nsteps - 1):
                self._t += self._dt
        self._invert()
        self._do_advection()
        self._do_friction()
        self._forward_timestep()

    # ------------------------------------------------------------- properties
    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, dt):
        self._dt = dt
        self._ablevel = 0

    @property
    def t(self):
        return self._t

    @property
    def tc(self):
        return self._tc

    @property
    def ablevel(self):
        return self._ablevel

    @property
    def rek(self):
        return self._rek

    @rek.setter
    def rek(self, rek):
        if self._pre_init("rek", rek):
            return
        self._rek = float(rek)
        self._upload_const("rek")

    @property
    def kk(self):
        return self._kk_v

    @kk.setter
    def kk(self, k):
        k = _typed(k, DTYPE_real, 1, "k")
        if self._pre_init("kk", k):
            return
        self._kk_v = np.array(k)
        self._ik_v = 1j * self._kk_v
        self._k2l2_v = self._kk_v[np.newaxis, :] ** 2 + self._ll_v[:, np.newaxis] ** 2
        self._upload_const("kk")
        self._upload_const("k2l2")

    @property
    def ll(self):
        return self._ll_v

    @ll.setter
    def ll(self, l):
        l = _typed(l, DTYPE_real, 1, "l")
        if self._pre_init("ll", l):
            return
        self._ll_v = np.array(l)
        self._il_v = 1j * self._ll_v
        self._k2l2_v = self._kk_v[np.newaxis, :] ** 2 + self._ll_v[:, np.newaxis] ** 2
        self._upload_const("ll")
        self._upload_const("k2l2")

    @property
    def _ik(self):
        return self._ik_v

    @property
    def _il(self):
        return self._il_v

    @property
    def _k2l2(self):
        self._const_exposed.add("k2l2")
        return self._k2l2_v

    @_k2l2.setter
    def _k2l2(self, v):
        v = _typed(v, DTYPE_real, 2, "_k2l2")
        self._k2l2_v[:] = v
        self._upload_const("k2l2")

    @property
    def a(self):
        self._const_exposed.add("a")
        return self._a_v

    @a.setter
    def a(self, b):
        b = _typed(b, DTYPE_real, 4, "b")
        if self._pre_init("a", b):
            return
        self._a_v[:] = b.astype(DTYPE_com)
        self._upload_const("a")

    @property
    def Ubg(self):
        return self._Ubg_v

    @Ubg.setter
    def Ubg(self, U):
        U = _typed(U, DTYPE_real, 1, "Ubg")
        if self._pre_init("Ubg", U):
            return
        self._Ubg_v = np.array(U)
        self._upload_const("Ubg")

    @property
    def Qy(self):
        return self._Qy_v

    @Qy.setter
    def Qy(self, Qy):
        Qy = _typed(Qy, DTYPE_real, 1, "Qy")
        if self._pre_init("Qy", Qy):
            return
        self._Qy_v = np.array(Qy)
        self._ikQy_v = 1j * (self._kk_v[np.newaxis, :] * self._Qy_v[:, np.newaxis])
        self._upload_const("Qy")

    @property
    def _ikQy(self):
        return self._ikQy_v

    @property
    def filtr(self):
        self._const_exposed.add("filtr")
        return self._filtr_v

    @filtr.setter
    def filtr(self, f):
        if self._pre_init("filtr", f):
            return
        self._filtr_v[:] = np.asarray(f, dtype=DTYPE_real)
        self._upload_const("filtr")

    @property
    def q(self):
        return self._get_field("q")

    @q.setter
    def q(self, b):
        b = _typed(b, DTYPE_real, 3, "b")
        if b.shape != (self.nz, self.ny, self.nx):
            raise ValueError(f"q: expected shape {(self.nz, self.ny, self.nx)}, got {b.shape}")
        self._flush()
        self._bucket.put("q", b)
        host = self._host.setdefault("q", np.empty((self.nz, self.ny, self.nx)))
        host[:] = b
        self._valid.add("q")
        _lib.check(_lib.lib.qgk_q_to_qh(self._bucket.h))
        self._wrote("q_set")

    @property
    def qh(self):
        return self._get_field("qh")

    @qh.setter
    def qh(self, b):
        b = _typed(b, DTYPE_com, 3, "b")
        if b.shape != (self.nz, self.nl, self.nk):
            raise ValueError(f"qh: expected shape {(self.nz, self.nl, self.nk)}, got {b.shape}")
        self._flush()
        self._bucket.put("qh", b)
        host = self._host.setdefault("qh", np.empty((self.nz, self.nl, self.nk), DTYPE_com))
        host[:] = b
        self._valid.add("qh")
        _lib.check(_lib.lib.qgk_qh_to_q(self._bucket.h))
        self._wrote("qh_set")

    @property
    def ufull(self):
        return self._get_field("u") + self._Ubg_v[:, np.newaxis, np.newaxis]

    @property
    def vfull(self):
        return self._get_field("v")


def _field_property(name):
    return property(lambda self: self._get_field(name),
                    doc=f"`{name}` (read-only alias of the device state, kernel.pyx:592-637)")


for _name in ("ph", "u", "v", "uh", "vh", "uq", "vq", "uqh", "vqh", "dqhdt", "dqhdt_p",
              "dqhdt_pp", "dqh", "duh", "dvh"):
    setattr(PseudoSpectralKernel, _name, _field_property(_name))


# module-level AB helpers (kernel.pyx:642-663)
def tendency_forward_euler(dt, dqdt):
    """Compute tendency using forward euler timestepping."""
    return dt * dqdt


def tendency_ab2(dt, dqdt, dqdt_p):
    """Compute tendency using Adams Bashforth 2nd order timestepping."""
    DT1 = 1.5 * dt
    DT2 = -0.5 * dt
    return DT1 * dqdt + DT2 * dqdt_p


def tendency_ab3(dt, dqdt, dqdt_p, dqdt_pp):
    """Compute tendency using Adams Bashforth 3rd order timestepping."""
    DT1 = 23 / 12. * dt
    DT2 = -16 / 12. * dt
    DT3 = 5 / 12. * dt
    return DT1 * dqdt + DT2 * dqdt_p + DT3 * dqdt_pp
