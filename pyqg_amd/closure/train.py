"""Training and offline evaluation of the closure network (libqgcnn kernels)."""
import math
import time

import numpy as np
import torch

from .cnn import CNN
from .param import Scales


def _normalized(x, y, scales, idx, net):
    dev = x.device
    xs = torch.as_tensor(scales.in_scale, dtype=torch.float32, device=dev)
    ys = torch.as_tensor(scales.out_scale, dtype=torch.float32, device=dev)
    return net.pad_input(x[idx] / xs), (y[idx] / ys).contiguous()


def train(net: CNN, x, y, scales: Scales, epochs=4, batch=64, lr=1e-3, lr_min=1e-5, seed=0,
          log=None, clip_factor=4.0, ema=0.998, ema_from=0.5, spec_weight=0.0):
    """Adam on the MSE of normalised forcing, cosine learning-rate decay.
    x (N, H, W, 6) and y (N, H, W, 2) FP32 tensors on the GPU.

    After 50 steps the global gradient norm is clipped at clip_factor times
    its running mean: without it one outsized gradient at the peak learning
    rate could switch off the ReLUs and the loss return to 1 for good (one
    seed in three at the default settings).

    With ema > 0 the network ends with an exponential moving average of its
    weights over the steps after the fraction ema_from of training (Polyak
    averaging): the closed model's long-run statistics are sensitive to the
    last steps' weight noise, which the average smooths."""
    n = x.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    steps = epochs * max(n // batch, 1)
    it = 0
    history = []
    gn_mean = 0.0
    avg = None
    t0 = time.perf_counter()
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g)
        for s in range(max(n // batch, 1)):
            idx = perm[s * batch:(s + 1) * batch].to(x.device)
            xb, yb = _normalized(x, y, scales, idx, net)
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * it / steps))
            clip = clip_factor * gn_mean if (clip_factor and it >= 50) else None
            loss = net.train_step(xb, yb, lr=cur, clip=clip, spec_weight=spec_weight)
            if ema and it >= int(ema_from * steps):
                params = list(net.w32) + list(net.b32)
                if avg is None:
                    avg = [p.clone() for p in params]
                else:
                    torch._foreach_mul_(avg, ema)
                    torch._foreach_add_(avg, params, alpha=1.0 - ema)
            gn = float(net.last_grad_norm)
            used = min(gn, clip) if clip is not None else gn
            gn_mean = used if it == 0 else 0.98 * gn_mean + 0.02 * used
            it += 1
            if it % 50 == 0 or it == steps:
                history.append((it, float(loss)))
                if log:
                    log(f"step {it}/{steps} loss {float(loss):.4f} grad norm {gn:.3g}")
    if avg is not None:
        nl = net.nlayers
        for l in range(nl):
            net.w32[l].copy_(avg[l])
            net.b32[l].copy_(avg[nl + l])
        net.refresh_bf16()
    torch.cuda.synchronize()
    return {"steps": it, "seconds": time.perf_counter() - t0, "history": history,
            "ema": ema if avg is not None else 0.0}


@torch.no_grad()
def evaluate(net: CNN, x, y, scales: Scales, batch=128):
    """Offline skill on held-out pairs: normalised MSE of the network and of
    the no-closure prediction (zero forcing), and R^2 = 1 - their ratio."""
    se, ss = 0.0, 0.0
    for s in range(0, x.shape[0], batch):
        idx = torch.arange(s, min(s + batch, x.shape[0]), device=x.device)
        xb, yb = _normalized(x, y, scales, idx, net)
        out = net.forward(xb)[..., :yb.shape[-1]]
        se += float(((out - yb) ** 2).sum())
        ss += float((yb ** 2).sum())
    return {"mse_net": se / ss, "mse_zero": 1.0, "r2": 1.0 - se / ss}


@torch.no_grad()
def amplitude(net, x, y, scales: Scales, batch=128):
    """Per output channel (layer) on held-out pairs: mean square of the target,
    of the prediction, and their mean product (normalised units).  An MSE
    fit shrinks the prediction towards the conditional mean: its mean square
    falls short of the target's by the unexplained variance.  `net` may be a
    list of networks with one input layout: the prediction is their mean."""
    nets = list(net) if isinstance(net, (list, tuple)) else [net]
    net = nets[0]
    nout = y.shape[-1]
    yy = torch.zeros(nout, dtype=torch.float64, device=x.device)
    pp = torch.zeros_like(yy)
    yp = torch.zeros_like(yy)
    n = 0
    for s in range(0, x.shape[0], batch):
        idx = torch.arange(s, min(s + batch, x.shape[0]), device=x.device)
        xb, yb = _normalized(x, y, scales, idx, net)
        out = sum(n.forward(xb)[..., :nout].double() for n in nets) / len(nets)
        yb = yb.double()
        yy += (yb ** 2).sum(dim=(0, 1, 2))
        pp += (out ** 2).sum(dim=(0, 1, 2))
        yp += (out * yb).sum(dim=(0, 1, 2))
        n += yb.shape[0] * yb.shape[1] * yb.shape[2]
    return {"target_ms": (yy / n).tolist(), "pred_ms": (pp / n).tolist(),
            "cross": (yp / n).tolist()}


@torch.no_grad()
def residual_statistics(net: CNN, x, y, scales: Scales, lag, batch=128):
    """The part of the held-out forcing the network leaves out, r = y - net(x)
    in the forcing's physical units: per layer, the rms of its unnormalised 2-D
    real transform (the units of the model's dqh) over samples, shape
    (nz, ny, nx/2 + 1), and the correlation of r between samples `lag` apart
    (the same member one sampling interval later when the pairs are stored
    snapshot by snapshot, `lag` members per snapshot)."""
    nout = y.shape[-1]
    osc = torch.as_tensor(scales.out_scale[:nout], dtype=torch.float64, device=x.device)
    res = []
    for s in range(0, x.shape[0], batch):
        idx = torch.arange(s, min(s + batch, x.shape[0]), device=x.device)
        xb, yb = _normalized(x, y, scales, idx, net)
        out = net.forward(xb)[..., :nout]
        res.append(((yb - out).double() * osc).float())
    r = torch.cat(res)                                  # (N, ny, nx, nz)
    spec = torch.zeros((nout, r.shape[1], r.shape[2] // 2 + 1), dtype=torch.float64,
                       device=x.device)
    for s in range(0, r.shape[0], batch):
        rh = torch.fft.rfft2(r[s:s + batch].permute(0, 3, 1, 2).double())
        spec += (rh.real ** 2 + rh.imag ** 2).sum(dim=0)
    amp = torch.sqrt(spec / r.shape[0]).cpu().numpy()
    a, b = r[:-lag].double(), r[lag:].double()
    rho = ((a * b).sum(dim=(0, 1, 2)) / torch.sqrt((a * a).sum(dim=(0, 1, 2)) *
                                                   (b * b).sum(dim=(0, 1, 2)))).cpu().numpy()
    return amp, rho
