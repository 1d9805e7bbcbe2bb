"""CPU oracle of pyqg's built-in closures (pyqg/parameterizations.py:132-289).

TEST INFRASTRUCTURE ONLY: the checker of the device closures; the package
never imports it.  Restated for the oracle models of qg_oracle.py (their
fft/ifft are numpy's, as in the reference's numpy build); `m.ik`, `m.il` of
the reference models are the 2-D grids 1j*k, 1j*l (model.py:464-470).

Pinned in tests/test_oracle.py against the reference's own outputs on the
same states (tests/golden/parameterizations.npz) and against the addition
and scaling identities of pyqg/tests/test_parameterizations.py:39-52.
"""
import numpy as np


def _ik(m):
    return 1j * m.k


def _il(m):
    return 1j * m.l


def smagorinsky(m, constant=0.1, just_viscosity=False):
    """parameterizations.py:160-188"""
    ik, il = _ik(m), _il(m)
    uh = m.fft(m.u)
    vh = m.fft(m.v)
    Sxx = m.ifft(uh * ik)
    Syy = m.ifft(vh * il)
    Sxy = 0.5 * m.ifft(uh * il + vh * ik)
    nu = (constant * m.dx) ** 2 * np.sqrt(2 * (Sxx ** 2 + Syy ** 2 + 2 * Sxy ** 2))
    if just_viscosity:
        return nu
    nu_Sxxh = m.fft(nu * Sxx)
    nu_Sxyh = m.fft(nu * Sxy)
    nu_Syyh = m.fft(nu * Syy)
    du = 2 * (m.ifft(nu_Sxxh * ik) + m.ifft(nu_Sxyh * il))
    dv = 2 * (m.ifft(nu_Sxyh * ik) + m.ifft(nu_Syyh * il))
    return du, dv


def backscatter_biharmonic(m, smag_constant=0.08, back_constant=0.99, eps=1e-32):
    """parameterizations.py:224-236"""
    ik, il = _ik(m), _il(m)
    lap = ik ** 2 + il ** 2
    psi = m.ifft(m.ph)
    lap_lap_psi = m.ifft(lap ** 2 * m.ph)
    dissipation = -m.ifft(lap * m.fft(lap_lap_psi * m.dx ** 2
                                      * smagorinsky(m, smag_constant, just_viscosity=True)))
    backscatter = -back_constant * lap_lap_psi * (
        (np.sum(m.Hi * np.mean(psi * dissipation, axis=(-1, -2))))
        / (np.sum(m.Hi * np.mean(psi * lap_lap_psi, axis=(-1, -2))) + eps))
    return dissipation + backscatter


def zanna_bolton_2020(m, constant=-46761284):
    """parameterizations.py:262-285"""
    ik, il = _ik(m), _il(m)
    uh = m.fft(m.u)
    vh = m.fft(m.v)
    vx = m.ifft(vh * ik)
    vy = m.ifft(vh * il)
    ux = m.ifft(uh * ik)
    uy = m.ifft(uh * il)
    rel_vort = vx - uy
    shearing = vx + uy
    stretching = ux - vy
    rv_stretch = m.fft(rel_vort * stretching)
    rv_shear = m.fft(rel_vort * shearing)
    sum_sqs = m.fft(rel_vort ** 2 + shearing ** 2 + stretching ** 2) / 2.0
    du = constant * m.ifft(ik * (sum_sqs - rv_shear) + il * rv_stretch)
    dv = constant * m.ifft(il * (sum_sqs + rv_shear) + ik * rv_stretch)
    return du, dv
