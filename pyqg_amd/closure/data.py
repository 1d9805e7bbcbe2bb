"""Training pairs for the learned closure from high-resolution runs of this solver.

A high-resolution two-layer QG ensemble (full state) is spun up and sampled;
at every sample time libqgk computes, for all members at once,

  coarse state   qbar = C(q)  (and ubar, vbar from the coarse inversion)
  forcing        S    = C(T_hi(q)) - T_lo(qbar)

where C is spectral truncation to the coarse grid times a Gaussian filter
exp(-(kappa Delta)^2 / 24), Delta the coarse grid spacing, and T the model
tendency of _invert + _do_advection (pyqg/kernel.pyx:312-395; friction and the
beta term are linear and cancel in S).  The pairs (qbar, ubar, vbar) -> S stay
on the GPU as FP32 NHWC tensors.  pyqg removed its own subgrid_forcing
(pyqg/model.py:397-402), so C is this project's definition.
"""
import numpy as np
import torch

from .. import _lib
from .. import constants as C
from ..ensemble import Ensemble

QG_EDDY = dict(L=1e6, rd=15000.0, delta=0.25, U1=0.025, U2=0.0, beta=1.5e-11, rek=5.787e-7,
               H1=500.0, filterfac=23.6)


def coarse_filter(n_lo, L, kind="gaussian"):
    """Filter on the coarse rfft2 grid applied after the spectral truncation:
    "gaussian", exp(-(kappa Delta)^2 / 24) with Delta = L / n_lo, or "sharp",
    the truncation alone (every resolved mode kept whole)."""
    g = C.grid(n_lo, n_lo, L, L)
    if kind == "sharp":
        return np.ones_like(g["wv2"])
    if kind != "gaussian":
        raise ValueError(f"unknown coarse filter {kind!r}")
    delta = L / n_lo
    return np.exp(-g["wv2"] * delta ** 2 / 24.0)


def _nhwc(b, field):
    """(M, nz, n, n) float64 device field -> (M, n, n, nz) float32 torch tensor."""
    t = torch.empty((b.M, b.nz, b.ny, b.nx), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(_lib.lib.qgk_get_device(b.h, _lib.FIELD_ID[field], ctypes_ptr(t)))
    return t.permute(0, 2, 3, 1).float().contiguous()


def ctypes_ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


class PairGenerator:
    """High-resolution ensemble + its coarse-graining onto a coarse grid."""

    def __init__(self, nmembers=64, n_hi=256, n_lo=64, dt=3600., seed=0, phys=None,
                 filter_kind="gaussian"):
        self.phys = dict(QG_EDDY, **(phys or {}))
        self.n_hi, self.n_lo, self.dt = n_hi, n_lo, dt
        self.hi = Ensemble.qg(np.full(nmembers, n_hi), dt=dt, full_state=True, **self.phys)
        self.lo = Ensemble.qg(np.full(nmembers, n_lo), dt=dt, full_state=True, q_param=True,
                              **self.phys)
        rng = np.random.default_rng(seed)
        self.hi.set_q([1e-7 * rng.standard_normal((2, n_hi, n_hi)) for _ in range(nmembers)])
        self.G = np.ascontiguousarray(coarse_filter(n_lo, self.phys["L"], filter_kind))

    def spin_up(self, nsteps, chunk=2000):
        while nsteps > 0:
            s = min(chunk, nsteps)
            self.hi.step(s)
            nsteps -= s

    def sample(self):
        """Pairs at the current state: inputs (M, n, n, 6) = (q1, q2, u1, u2, v1,
        v2) and targets (M, n, n, 2) = (S1, S2), FP32 on the GPU."""
        hb, lb = self.hi.buckets[0], self.lo.buckets[0]
        _lib.check(_lib.lib.qgk_subgrid_forcing(hb.h, lb.h, _lib.ptr(self.G)))
        x = torch.cat([_nhwc(lb, "q"), _nhwc(lb, "u"), _nhwc(lb, "v")], dim=-1)
        y = _nhwc(lb, "dq")
        return x, y

    def hi_ph(self):
        """The high-resolution streamfunction of every member at the last
        sample (host array), for reference spectra."""
        return self.hi.buckets[0].get("ph")

    def generate(self, nsamples, every=24):
        xs, ys = [], []
        n = 0
        while n < nsamples:
            self.hi.step(every)
            x, y = self.sample()
            xs.append(x)
            ys.append(y)
            n += x.shape[0]
        return torch.cat(xs)[:nsamples], torch.cat(ys)[:nsamples]
