"""Isotropic kinetic-energy spectra (the closure's evaluation metric).

Follows the definitions of pyqg's KEspec diagnostic (pyqg/model.py:657-663:
wv2 |ph|^2 / M^2 per layer) and of diagnostic_tools.calc_ispec
(pyqg/diagnostic_tools.py:51-126: sum over annuli of width dk, counting the
half-plane modes twice except kx = 0 and kx = nx/2), weighted by layer
thickness.
"""
import numpy as np


def ke_density(ph, wv2, Hi):
    """Layer-thickness weighted KE density on the rfft2 grid: (nl, nk)."""
    M = ph.shape[-2] * (2 * (ph.shape[-1] - 1))
    H = float(np.sum(Hi))
    e = wv2[np.newaxis] * np.abs(ph) ** 2 / M ** 2
    return np.tensordot(np.asarray(Hi, dtype=float) / H, e, axes=(0, 0))


def isotropic(field, kk, ll, kmax=None):
    """Sum an (nl, nk) rfft2-grid quantity over annuli of width dk = kk[1]."""
    dk = kk[1] - kk[0]
    k, l = np.meshgrid(kk, ll)
    kappa = np.sqrt(k ** 2 + l ** 2)
    w = np.full(field.shape, 2.0)
    w[:, 0] = 1.0
    w[:, -1] = 1.0
    kmax = kmax if kmax is not None else kk.max()
    edges = np.arange(0.5 * dk, kmax + dk, dk)
    centers = 0.5 * (edges[1:] + edges[:-1])
    spec = np.zeros(len(centers))
    idx = np.digitize(kappa.ravel(), edges) - 1
    ok = (idx >= 0) & (idx < len(centers))
    np.add.at(spec, idx[ok], (w * field).ravel()[ok])
    return centers, spec


def ke_spectrum(ph, kk, ll, wv2, Hi, kmax=None):
    """Isotropic KE spectrum of one state: (kappa, E(kappa))."""
    return isotropic(ke_density(ph, wv2, Hi), kk, ll, kmax)


def spectrum_error(e_test, e_ref):
    """Relative L2 difference of two spectra on their common bins."""
    n = min(len(e_test), len(e_ref))
    a, b = np.asarray(e_test[:n]), np.asarray(e_ref[:n])
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
