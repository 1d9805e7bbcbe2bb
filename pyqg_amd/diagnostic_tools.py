"""Spectral utilities and model comparisons (pyqg/diagnostic_tools.py:1-287).

They run on the host over diagnostics the device has already time-averaged
(a (nl, nk) or (nz, nl, nk) array per model), so nothing here touches the
per-step path.  `grid(...)` gives the wavenumber attributes these functions
read when there is no model object at hand (e.g. an ensemble's mean spectrum).
"""
import types

import numpy as np
from numpy import pi

from . import constants as C


def grid(nx, L, ny=None, W=None):
    """The attributes calc_ispec reads (kk, ll, dk, dl, wv, M) of an nx x ny grid."""
    ny = nx if ny is None else ny
    W = L if W is None else W
    g = C.grid(nx, ny, L, W)
    return types.SimpleNamespace(kk=g["kk"], ll=g["ll"], dk=g["dk"], dl=g["dl"], wv=g["wv"],
                                 wv2=g["wv2"], M=g["M"], nx=nx, ny=ny)


def spec_var(model, ph):
    """Variance of a field from its rfft2 coefficients ph (diagnostic_tools.py:6-28):
    the half-plane sum counts the k = 0 and k = nx/2 columns once, the others twice."""
    var_dens = 2. * np.abs(ph) ** 2 / model.M ** 2
    var_dens[..., 0] /= 2
    var_dens[..., -1] /= 2
    return var_dens.sum(axis=(-1, -2))


def spec_sum(ph2):
    """Total of a real half-plane spectral density (diagnostic_tools.py:31-52)."""
    ph2 = 2. * ph2
    ph2[..., 0] = ph2[..., 0] / 2.
    ph2[..., -1] = ph2[..., -1] / 2.
    return ph2.sum(axis=(-1, -2))


def calc_ispec(model, _var_dens, averaging=True, truncate=True, nd_wavenumber=False,
               nfactor=1):
    """Isotropic spectrum of a half-plane spectral density
    (diagnostic_tools.py:54-126): bins of width nfactor * |(dk, dl)| from 0 to
    the inner (truncate) or outer circle; with averaging the bin mean times the
    ring circumference in wavenumber cells, else the bin sum over the width.
    Returns (bin centres, spectrum)."""
    var_dens = np.copy(_var_dens)
    var_dens[..., 0] /= 2
    var_dens[..., -1] /= 2
    ll_max = np.abs(model.ll).max()
    kk_max = np.abs(model.kk).max()
    kmax = np.minimum(ll_max, kk_max) if truncate else np.sqrt(ll_max ** 2 + kk_max ** 2)
    kmin = 0
    dkr = np.sqrt(model.dk ** 2 + model.dl ** 2) * nfactor
    kr = np.arange(kmin, kmax, dkr)
    phr = np.zeros(kr.size)
    for i in range(kr.size):
        if i == kr.size - 1:
            sel = (model.wv >= kr[i]) & (model.wv <= kr[i] + dkr)
        else:
            sel = (model.wv >= kr[i]) & (model.wv < kr[i + 1])
        if averaging:
            phr[i] = var_dens[sel].mean() * (kr[i] + dkr / 2) * pi / (model.dk * model.dl)
        else:
            phr[i] = var_dens[sel].sum() / dkr
        phr[i] *= 2
    kr = kr + dkr / 2
    if nd_wavenumber:
        # The reference divides by kmin, which is 0 there (diagnostic_tools.py:
        # 118-121), giving inf; here the wavenumbers are made nondimensional
        # with the fundamental wavenumber of the grid instead.
        kfund = min(model.dk, model.dl)
        kr = kr / kfund
        phr = phr * kfund
    return kr, phr


def diagnostic_differences(m1, m2, reduction='rmse', instantaneous=False):
    """Distance between the diagnostics of two models, possibly of different
    resolution (diagnostic_tools.py:128-229): (l, k) diagnostics are compared
    as isotropic spectra over the wavenumbers both resolve, per layer for
    layered ones, with their `sums_with` partners added; entries flagged
    skip_comparison are left out.  Returns {name[layer]: distance}."""
    diffs = {}
    kr1, _ = calc_ispec(m1, m1.diagnostics['KEspec']['function'](m1)[0])
    kr2, _ = calc_ispec(m2, m2.diagnostics['KEspec']['function'](m2)[0])
    nkr = min(len(kr1), len(kr2))

    def normalized(model, name, layer=None):
        attrs = model.diagnostics[name]
        diag = attrs['function'](model) if instantaneous else model.get_diagnostic(name)
        diag = np.array(diag, dtype=np.float64)
        for other in attrs.get('sums_with', []):
            diag = diag + (model.diagnostics[other]['function'](model) if instantaneous
                           else model.get_diagnostic(other))
        if layer is not None:
            diag = diag[layer]
        if tuple(attrs['dims'][-2:]) == ('l', 'k'):
            _, diag = calc_ispec(model, diag)
            diag = diag[:nkr]
        return diag

    for name, attrs in m1.diagnostics.items():
        if attrs.get('skip_comparison', False) or name not in m2.diagnostics:
            continue
        if attrs['dims'][0] == 'lev':
            layers = range(m1.nz)
        elif attrs['dims'][0] == 'lev_mid':
            layers = range(m1.nz - 1)
        else:
            layers = [None]
        for layer in layers:
            d1 = normalized(m1, name, layer)
            d2 = normalized(m2, name, layer)
            label = f"{name}{'' if layer is None else layer + 1}"
            diffs[label] = (np.sqrt(np.mean((d1 - d2) ** 2)) if reduction == 'rmse'
                            else reduction(d1, d2))
    return diffs


def diagnostic_similarities(model, target, baseline, **kw):
    """1 - d(model, target) / d(baseline, target) per diagnostic
    (diagnostic_tools.py:231-262): near 1 when the model is much closer to the
    target than the baseline is, negative when it is farther."""
    d1 = diagnostic_differences(model, target, **kw)
    d2 = diagnostic_differences(baseline, target, **kw)
    return dict((k, 1 - d1[k] / d2[k]) for k in d1.keys())
