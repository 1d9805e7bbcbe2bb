"""Host-side builders of the constants the kernel consumes.

Shared by the Model front-ends and by Ensemble.  Each follows the reference's
definition: grid and wavenumbers (pyqg/model.py:446-482), spectral filter
(model.py:496-503), two-layer background and inversion (qg_model.py:110-161),
N-layer stretching matrix and inversion (layered_model.py:136-229).
"""
import numpy as np
from numpy import pi


def grid(nx, ny, L, W):
    """Wavenumbers of an rfft2 grid.  ll is built from nx, as pyqg does
    (model.py:461-462), so only nx == ny is meaningful."""
    nk = nx // 2 + 1
    g = {}
    g["x"], g["y"] = np.meshgrid(np.arange(0.5, nx, 1.) / nx * L, np.arange(0.5, ny, 1.) / ny * W)
    g["dk"] = 2. * pi / L
    g["dl"] = 2. * pi / W
    half = nx / 2
    g["ll"] = g["dl"] * np.concatenate([np.arange(0., half), np.arange(-half, 0.)])
    g["kk"] = g["dk"] * np.arange(0., nk)
    g["k"], g["l"] = np.meshgrid(g["kk"], g["ll"])
    g["dx"] = L / nx
    g["dy"] = W / ny
    g["M"] = nx * ny
    g["wv2"] = g["k"] ** 2 + g["l"] ** 2
    g["wv"] = np.sqrt(g["wv2"])
    nz = g["wv2"] != 0.
    g["wv2i"] = np.where(nz, 1. / np.where(nz, g["wv2"], 1.), 0.)
    # the kernel's _k2l2 (kernel.pyx:543-546)
    g["k2l2"] = g["kk"][np.newaxis, :] ** 2 + g["ll"][:, np.newaxis] ** 2
    return g


def spectral_filter(k, l, dx, dy, filterfac):
    """Exponential cutoff filter, 1 below 0.65 pi (model.py:496-503)."""
    cphi = 0.65 * pi
    wvx = np.sqrt((k * dx) ** 2. + (l * dy) ** 2.)
    return np.where(wvx <= cphi, 1., np.exp(-filterfac * (wvx - cphi) ** 4.))


def qg_background(beta, rd, delta, U1, U2):
    """F1, F2, Qy, Ubg of the two-layer model (qg_model.py:119-126)."""
    F1 = rd ** -2 / (1. + delta)
    F2 = delta * F1
    Qy = np.array([beta + F1 * (U1 - U2), beta - F2 * (U1 - U2)])
    return F1, F2, Qy, np.array([U1, U2], dtype=np.float64)


def qg_inversion(wv2, F1, F2):
    """Closed-form inverse of (S - kappa^2 I) for two layers, zero at kappa = 0
    (qg_model.py:146-161)."""
    det = wv2 * (wv2 + F1 + F2)
    ok = det != 0.
    di = np.where(ok, 1. / np.where(ok, det, 1.), 0.)
    a = np.empty((2, 2) + wv2.shape)
    a[0, 0] = -(wv2 + F2) * di
    a[0, 1] = -F1 * di
    a[1, 0] = -F2 * di
    a[1, 1] = -(wv2 + F1) * di
    return a


def layered_stretching(nz, Hi, rhoi, g, f2=None, rd=None, delta=None):
    """Stretching matrix S and reduced gravities (layered_model.py:136-193)."""
    S = np.zeros((nz, nz))
    gpi = None
    if nz == 2 and rd and delta:
        F1 = rd ** -2 / (1. + delta)
        F2 = delta * F1
        S[0, 0], S[0, 1] = -F1, F1
        S[1, 0], S[1, 1] = F2, -F2
        return S, gpi
    gpi = g * (rhoi[1:] - rhoi[:-1]) / rhoi[:-1]
    assert gpi.size == nz - 1, "Invalid size of gpi"
    assert np.all(gpi > 0.), "Buoyancy jump has negative sign!"
    c = f2 / Hi
    for i in range(nz):
        up = c[i] / gpi[i - 1] if i > 0 else 0.
        dn = c[i] / gpi[i] if i < nz - 1 else 0.
        if i > 0:
            S[i, i - 1] = up
        if i < nz - 1:
            S[i, i + 1] = dn
        S[i, i] = -(dn + up) if 0 < i < nz - 1 else -(dn if i == 0 else up)
    return S, gpi


def layered_inversion(S, wv2):
    """a = (S - kappa^2 I)^-1 at every wavenumber, zero at kappa = 0
    (layered_model.py:211-229); returned as [nz][nz][nl][nk]."""
    nz = S.shape[0]
    M = S[np.newaxis, np.newaxis] - np.eye(nz) * wv2[..., np.newaxis, np.newaxis]
    a = np.zeros(wv2.shape + (nz, nz))
    sel = wv2 != 0.
    sel[0, 0] = False
    a[sel] = np.linalg.inv(M[sel])
    a[~np.isfinite(a)] = 0.
    return np.ascontiguousarray(np.transpose(a, (2, 3, 0, 1)))


def exponential_stratification(nz=10, Htot=4000.0, hscale=800.0, uscale=1000.0,
                               rd1=20000.0, f=0.0001236812857687059, g=9.81):
    """Layer thicknesses H_k ~ exp(-z_k / hscale) (z_k the layer mid-depths,
    summing to Htot), U_k = 0.05 exp(-z_k / uscale), and densities whose
    reduced gravities are scaled so that the first baroclinic deformation
    radius of the stretching matrix is rd1.  Returns (H, rho, U)."""
    H = np.full(nz, Htot / nz)
    for _ in range(200):
        z = np.cumsum(H) - 0.5 * H
        Hn = np.exp(-z / hscale)
        Hn *= Htot / Hn.sum()
        if np.allclose(Hn, H, rtol=0, atol=1e-12):
            break
        H = Hn
    z = np.cumsum(H) - 0.5 * H
    U = 0.05 * np.exp(-z / uscale)
    gp = np.full(nz - 1, 0.01)

    def radius(gp):
        S = np.zeros((nz, nz))
        f2 = f * f
        for i in range(nz):
            if i > 0:
                S[i, i - 1] = f2 / H[i] / gp[i - 1]
                S[i, i] -= f2 / H[i] / gp[i - 1]
            if i < nz - 1:
                S[i, i + 1] = f2 / H[i] / gp[i]
                S[i, i] -= f2 / H[i] / gp[i]
        ev = np.sort(np.linalg.eigvals(-S).real)
        return 1.0 / np.sqrt(ev[1])

    gp = gp * (rd1 / radius(gp)) ** 2
    rho = np.empty(nz)
    rho[0] = 1025.0
    for i in range(nz - 1):
        rho[i + 1] = rho[i] * (1.0 + gp[i] / g)
    return H, rho, U


def vertical_modes(S, Hi):
    """Deformation wavenumbers kdi2 (ascending eigenvalues of -S) and the
    pressure modes normalised so that sum_k Hi_k pmodes_k^2 = H
    (model.py:286-305).  Returns (kdi2, pmodes)."""
    evals, evecs = np.linalg.eig(-S)
    asort = evals.argsort()
    kdi2 = evals[asort]
    pmodes = evecs[:, asort]
    H = np.sum(Hi)
    Ai = (H / (Hi[:, np.newaxis] * (pmodes ** 2)).sum(axis=0)) ** 0.5
    return kdi2, Ai[np.newaxis, :] * pmodes


def ab_weights(ablevel, dt):
    """Adams-Bashforth weights and the next ablevel (kernel.pyx:482-500)."""
    if ablevel == 0:
        return (dt, 0.0, 0.0), 1
    if ablevel == 1:
        return (1.5 * dt, -0.5 * dt, 0.0), 2
    return (23. / 12. * dt, -16. / 12. * dt, 5. / 12. * dt), ablevel
