"""The packed x-mixed exchange of the fused register path, restated in NumPy
and checked against the oracle's fields (no GPU).

Between the column kernel (y transforms) and the row kernel (x transforms) a
row of an x-mixed field carries nx/2 slots instead of nx/2 + 1: the x
transform of real data makes bins 0 and nx/2 real, so one complex slot holds
both.  The column kernel transforms column 0 + i * (column nx/2) as one
sequence and separates the two spectra by their symmetry; going back it packs
the Hermitian parts of the two columns, the only parts the reference's c2r
uses (kernel.pyx:246-267: numpy's irfft drops the imaginary parts of bins 0
and nx/2).  These tests pin that algebra on the reference's own fields,
including the Nyquist column's anti-Hermitian parts (i k UQ and v = ik psi at
kx = nx/2) and the self-conjugate rows 0 and ny/2 of column 0.
"""
import numpy as np
import pytest

from conftest import QG_PHYS
from oracle.qg_oracle import OracleQG


def _hermitian_part(X):
    Xm = np.roll(X[..., ::-1], 1, axis=-1)  # X(-l)
    return 0.5 * (X + np.conj(Xm))


def pack_inverse(Xh, n):
    """y inverse transforms of columns 0..n/2-1 with slot 0 = col 0 + i col n/2
    (unnormalized), as the column kernel writes them."""
    N2 = n // 2
    X = np.fft.ifft(Xh, axis=-2) * n
    G0 = Xh[..., 0].copy()
    G0[..., 0] = G0[..., 0].real
    G0[..., N2] = G0[..., N2].real
    GN = _hermitian_part(Xh[..., N2])
    P = X[..., :N2].copy()
    P[..., 0] = np.fft.ifft(G0 + 1j * GN, axis=-1) * n
    return P


def row_c2r(P, n):
    N2 = n // 2
    H = np.zeros(P.shape[:-1] + (N2 + 1,), complex)
    H[..., :N2] = P
    H[..., 0] = P[..., 0].real
    H[..., N2] = P[..., 0].imag
    return np.fft.irfft(H, n=n, axis=-1) / n


def row_r2c_packed(x):
    N2 = x.shape[-1] // 2
    X = np.fft.rfft(x, axis=-1)
    P = X[..., :N2].copy()
    P[..., 0] = X[..., 0].real + 1j * X[..., N2].real
    return P


def col_forward_unpack(P):
    N2 = P.shape[-1]
    F = np.fft.fft(P, axis=-2)
    out = np.zeros(P.shape[:-1] + (N2 + 1,), complex)
    out[..., :N2] = F
    F0 = F[..., 0]
    Fm = np.roll(F0[..., ::-1], 1, axis=-1)
    out[..., 0] = 0.5 * (F0 + np.conj(Fm))
    out[..., N2] = (F0 - np.conj(Fm)) / 2j
    return out


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("filterfac,n", [(23.6, 64), (0.0, 64), (0.0, 48)])
def test_packed_exchange_reproduces_reference_fields(filterfac, n):
    phys = dict(QG_PHYS, filterfac=filterfac)
    o = OracleQG(nx=n, dt=3600., **phys)
    o.q = 1e-7 * np.random.default_rng(40).standard_normal((2, n, n))
    for step in range(3):
        o._invert()
        o._do_advection()
        kk, ll = np.asarray(o.kk), np.asarray(o.ll)
        qh, ph = np.asarray(o.qh), np.asarray(o.ph)
        uh = -1j * ll[None, :, None] * ph
        vh = 1j * kk[None, None, :] * ph
        q = row_c2r(pack_inverse(qh, n), n)
        u = row_c2r(pack_inverse(uh, n), n)
        v = row_c2r(pack_inverse(vh, n), n)
        assert _rel(q, np.asarray(o.q)) < 1e-14
        assert _rel(u, np.asarray(o.u)) < 1e-14
        assert _rel(v, np.asarray(o.v)) < 1e-14
        uq, vq = np.asarray(o.uq), np.asarray(o.vq)
        assert _rel(col_forward_unpack(row_r2c_packed(uq)), np.asarray(o.uqh)) < 1e-14
        assert _rel(col_forward_unpack(row_r2c_packed(vq)), np.asarray(o.vqh)) < 1e-14
        o._do_friction()
        o._forward_timestep()
    # after steps the Nyquist column carries anti-Hermitian parts (i k UQ)
    N2 = n // 2
    qN = np.asarray(o.qh)[..., N2]
    assert np.abs(qN - _hermitian_part(qN)).max() > 1e-3 * np.abs(qN).max()


def test_plain_packing_of_the_nyquist_column_would_be_wrong():
    """Packing column n/2 as it is (not its Hermitian part) sends v's
    anti-Hermitian Nyquist spectrum into the zonal-mean bin: the check above
    is sensitive to that mistake."""
    n = 64
    o = OracleQG(nx=n, dt=3600., **QG_PHYS)
    o.q = 1e-7 * np.random.default_rng(40).standard_normal((2, n, n))
    o._invert()
    kk = np.asarray(o.kk)
    vh = 1j * kk[None, None, :] * np.asarray(o.ph)
    N2 = n // 2
    X = np.fft.ifft(vh, axis=-2) * n
    P = X[..., :N2].copy()
    P[..., 0] = np.fft.ifft(vh[..., 0] + 1j * vh[..., N2], axis=-1) * n
    assert _rel(row_c2r(P, n), np.asarray(o.v)) > 1e-3
