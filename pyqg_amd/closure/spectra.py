"""KE-spectrum error of a coarse model against the high-resolution truth.

In pyqg's own terms (diagnostic_tools.diagnostic_differences /
diagnostic_similarities, pyqg/diagnostic_tools.py:128-262): the time-mean
KEspec diagnostic of each run (accumulated on the device, averaged over the
ensemble's members) is made isotropic per layer with calc_ispec, cut to the
wavenumbers both grids resolve, and compared by RMSE; the similarity of a
closed run is 1 - RMSE(closed, truth) / RMSE(plain, truth).
"""
import numpy as np

from ..diagnostic_tools import calc_ispec


def layer_spectra(kespec, g):
    """Isotropic spectrum of every layer of a (nz, nl, nk) KEspec: (kr, [nz][nkr])."""
    kr = None
    out = []
    for z in range(kespec.shape[0]):
        kr, s = calc_ispec(g, kespec[z])
        out.append(s)
    return kr, np.array(out)


def ke_errors(kespec, g, kespec_ref, g_ref):
    """{"KEspec1": rmse, "KEspec2": rmse, ...} on the common wavenumbers."""
    kr, s = layer_spectra(kespec, g)
    kr_ref, s_ref = layer_spectra(kespec_ref, g_ref)
    n = min(len(kr), len(kr_ref))
    return {f"KEspec{z + 1}": float(np.sqrt(np.mean((s[z, :n] - s_ref[z, :n]) ** 2)))
            for z in range(s.shape[0])}


def similarity(err_model, err_baseline):
    """1 - d(model, truth) / d(baseline, truth) per entry."""
    return {k: 1.0 - err_model[k] / err_baseline[k] for k in err_model}


def relative_errors(kespec, g, kespec_ref, g_ref):
    """Relative L2 difference of the per-layer isotropic spectra (common bins)."""
    kr, s = layer_spectra(kespec, g)
    kr_ref, s_ref = layer_spectra(kespec_ref, g_ref)
    n = min(len(kr), len(kr_ref))
    return {f"KEspec{z + 1}": float(np.linalg.norm(s[z, :n] - s_ref[z, :n])
                                    / np.linalg.norm(s_ref[z, :n]))
            for z in range(s.shape[0])}
