"""Training helpers of the learned closure that run on the CPU (torch only)."""
import numpy as np


def test_spectral_loss_gradient():
    """The log-spectral term's gradient: zero where the prediction already
    has the target's power at every wavenumber (here: the target itself, and
    the target shifted periodically, which keeps every power), zero on the
    padding channels, and the directional derivative of the loss it is the
    gradient of (central differences, float64)."""
    import torch
    from pyqg_amd.closure.cnn import spectral_loss_grad
    g = torch.Generator().manual_seed(0)
    tgt = torch.randn(3, 8, 8, 2, generator=g, dtype=torch.float64)
    out = torch.zeros(3, 8, 8, 4, dtype=torch.float64)
    out[..., :2] = tgt
    assert float(spectral_loss_grad(out, tgt).abs().max()) < 1e-12
    out[..., :2] = torch.roll(tgt, shifts=(3, 5), dims=(1, 2))
    assert float(spectral_loss_grad(out, tgt).abs().max()) < 1e-12
    out = torch.randn(3, 8, 8, 4, generator=g, dtype=torch.float64)
    grad = spectral_loss_grad(out, tgt)
    assert not grad[..., 2:].any() and grad[..., :2].abs().max() > 0

    def loss(o):
        po = torch.fft.rfft2(o[..., :2].permute(0, 3, 1, 2))
        pt = torch.fft.rfft2(tgt.permute(0, 3, 1, 2))
        eo = (po.abs() ** 2).mean(0)
        et = (pt.abs() ** 2).mean(0)
        eps = 1e-6 * et.mean()
        return float(((torch.log(eo + eps) - torch.log(et + eps)) ** 2).mean())

    d = torch.randn(out.shape, generator=g, dtype=torch.float64)
    h = 1e-6
    fd = (loss(out + h * d) - loss(out - h * d)) / (2 * h)
    np.testing.assert_allclose(float((grad * d).sum()), fd, rtol=1e-6)
