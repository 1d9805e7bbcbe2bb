"""Fully convolutional subgrid-forcing network on libqgcnn's kernels.

Every forward, backward and optimizer operation is a libqgcnn kernel (BF16
operands, FP32 accumulation on the matrix cores); torch tensors are only the
device-memory containers and provide the stream.  Activations are NHWC BF16
with channels padded to multiples of 16; weights [cout][25][cin].
"""
import ctypes
import math
import os

import numpy as np

from .. import _lib as _qgk  # noqa: F401  (one HIP runtime for torch and the kernels)
import torch  # noqa: E402

_HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_HERE, "_qgcnn.so")

DEFAULT_CHANNELS = (6, 128, 64, 32, 32, 32, 32, 32, 2)


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libqgcnn is not built ({LIB_PATH} missing): run "
                          "`python pyqg_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, F, Z = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
    sig = {
        "qgcnn_last_error": (ctypes.c_char_p, []),
        "qgcnn_conv_fwd": (I, [P, I, I, I, I, P, P, I, I, P, P, P, P]),
        "qgcnn_conv_wgrad": (I, [P, P, I, I, I, I, I, P, P]),
        "qgcnn_flip_transpose": (I, [P, P, I, I, P]),
        "qgcnn_bias_grad": (I, [P, Z, I, P, P]),
        "qgcnn_bias_grad_ws": (I, [P, Z, I, P, P, Z, P]),
        "qgcnn_mse": (I, [P, P, Z, I, I, F, P, P, P]),
        "qgcnn_f32_to_bf16": (I, [P, P, Z, P]),
        "qgcnn_bf16_to_f32": (I, [P, P, Z, P]),
        "qgcnn_adam": (I, [P, P, P, P, P, Z, F, F, F, F, I, F, P]),
        "qgcnn_wgrad_workspace_bytes": (Z, [I, I, I, I, I]),
        "qgcnn_conv_wgrad_ws": (I, [P, P, I, I, I, I, I, P, P, Z, P]),
        "qgcnn_pack_first": (I, [P, I, I, P, P]),
        "qgcnn_conv_first": (I, [P, I, I, I, I, P, P, I, I, P, P, P]),
        "qgcnn_pack_last": (I, [P, I, I, I, P, P]),
        "qgcnn_conv_last": (I, [P, I, I, I, I, P, P, I, I, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
EXPORTS = ["qgcnn_last_error", "qgcnn_conv_fwd", "qgcnn_conv_wgrad", "qgcnn_flip_transpose",
           "qgcnn_bias_grad", "qgcnn_mse", "qgcnn_f32_to_bf16", "qgcnn_bf16_to_f32",
           "qgcnn_adam", "qgcnn_wgrad_workspace_bytes", "qgcnn_conv_wgrad_ws",
           "qgcnn_bias_grad_ws", "qgcnn_pack_first", "qgcnn_conv_first", "qgcnn_pack_last",
           "qgcnn_conv_last"]


def check(rc):
    if rc != 0:
        msg = lib.qgcnn_last_error().decode(errors="replace")
        raise (ValueError if rc == -1 else RuntimeError)(msg)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def pad16(c):
    return ((c + 15) // 16) * 16


# ---------------------------------------------------------------- primitives
def conv_fwd(x, w, bias, coutp, relu=True, mask=None, out_bf16=True, out_f32=False, y=None,
             y32=None, stream=None):
    """x: (B, H, 64, cinp) bf16; w: (coutp, 25, cinp) bf16; returns (y_bf16, y_f32).
    Outputs are allocated unless given (y, y32); stream: a hipStream_t handle
    (default: torch's current stream)."""
    B, H, W, cinp = x.shape
    if out_bf16 and y is None:
        y = torch.empty((B, H, W, coutp), dtype=torch.bfloat16, device=x.device)
    if out_f32 and y32 is None:
        y32 = torch.empty((B, H, W, coutp), dtype=torch.float32, device=x.device)
    y = y if out_bf16 else None
    y32 = y32 if out_f32 else None
    check(lib.qgcnn_conv_fwd(_p(x), B, H, W, cinp, _p(w), _p(bias), coutp, int(relu), _p(mask),
                             _p(y), _p(y32), stream if stream is not None else _stream()))
    return y, y32


def conv_first(x, w8, bias, coutp, relu=True, y=None, y32=None, out_bf16=True, out_f32=False,
               stream=None):
    """First layer with at most 8 real input channels: as conv_fwd with the taps
    packed into the matrix K dimension; w8 from pack_first."""
    B, H, W, cinp = x.shape
    if out_bf16 and y is None:
        y = torch.empty((B, H, W, coutp), dtype=torch.bfloat16, device=x.device)
    if out_f32 and y32 is None:
        y32 = torch.empty((B, H, W, coutp), dtype=torch.float32, device=x.device)
    y = y if out_bf16 else None
    y32 = y32 if out_f32 else None
    check(lib.qgcnn_conv_first(_p(x), B, H, W, cinp, _p(w8), _p(bias), coutp, int(relu), _p(y),
                               _p(y32), stream if stream is not None else _stream()))
    return y, y32


def pack_first(w, w8=None):
    coutp, _, cinp = w.shape
    if w8 is None:
        w8 = torch.empty((coutp, 28, 8), dtype=torch.bfloat16, device=w.device)
    check(lib.qgcnn_pack_first(_p(w), cinp, coutp, _p(w8), _stream()))
    return w8


def conv_last(x, wl, bias, coutp, cout, y=None, y32=None, out_bf16=False, out_f32=True,
              stream=None):
    """Last layer with 5 * cout <= 16 real outputs: as conv_fwd without
    activation, the horizontal taps as matrix rows; wl from pack_last."""
    B, H, W, cinp = x.shape
    if out_bf16 and y is None:
        y = torch.empty((B, H, W, coutp), dtype=torch.bfloat16, device=x.device)
    if out_f32 and y32 is None:
        y32 = torch.empty((B, H, W, coutp), dtype=torch.float32, device=x.device)
    y = y if out_bf16 else None
    y32 = y32 if out_f32 else None
    check(lib.qgcnn_conv_last(_p(x), B, H, W, cinp, _p(wl), _p(bias), coutp, cout, _p(y),
                              _p(y32), stream if stream is not None else _stream()))
    return y, y32


def pack_last(w, cout, wl=None):
    coutp, _, cinp = w.shape
    if wl is None:
        wl = torch.empty((16, 5, cinp), dtype=torch.bfloat16, device=w.device)
    check(lib.qgcnn_pack_last(_p(w), cinp, coutp, cout, _p(wl), _stream()))
    return wl


_WS = {}


def _workspace(nbytes, device):
    """A device buffer of at least nbytes, kept for later calls."""
    ws = _WS.get(device)
    if ws is None or ws.numel() < nbytes:
        ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
        _WS[device] = ws
    return ws


def conv_wgrad(dy, x, dw, blocked=True):
    """dw += weight gradient.  blocked: the matrix-core kernel with a workspace
    (deterministic); otherwise the atomic kernel."""
    B, H, W, coutp = dy.shape
    cinp = x.shape[-1]
    if blocked:
        n = lib.qgcnn_wgrad_workspace_bytes(B, H, W, cinp, coutp)
        if n == 0:
            raise ValueError(f"unsupported shape for the weight gradient: {tuple(dy.shape)}, "
                             f"cinp={cinp}")
        ws = _workspace(n, dy.device)
        check(lib.qgcnn_conv_wgrad_ws(_p(dy), _p(x), B, H, W, cinp, coutp, _p(dw), _p(ws),
                                      ws.numel(), _stream()))
        return
    check(lib.qgcnn_conv_wgrad(_p(dy), _p(x), B, H, W, cinp, coutp, _p(dw), _stream()))


def flip_transpose(w, wt):
    coutp, _, cinp = w.shape
    check(lib.qgcnn_flip_transpose(_p(w), _p(wt), cinp, coutp, _stream()))


def bias_grad(dy, db):
    """db += the bias gradient, summed in a fixed order (bit-reproducible)."""
    coutp = dy.shape[-1]
    ws = _workspace(256 * coutp * 4, dy.device)
    check(lib.qgcnn_bias_grad_ws(_p(dy), dy.numel() // coutp, coutp, _p(db), _p(ws),
                                 ws.numel() // 4, _stream()))


def mse(pred, target, grad=None, gscale=1.0):
    """pred (B,H,W,coutp) f32, target (B,H,W,nch) f32 -> loss tensor (1,) f32."""
    loss = torch.zeros(1, dtype=torch.float32, device=pred.device)
    npix = pred.numel() // pred.shape[-1]
    check(lib.qgcnn_mse(_p(pred), _p(target), npix, pred.shape[-1], target.shape[-1],
                        float(gscale), _p(grad), _p(loss), _stream()))
    return loss


def spectral_loss_grad(out, target):
    """Gradient with respect to the network output of the log-spectral
    distance between prediction and target: mean over channels and 2-D
    wavenumbers of (log E_pred - log E_target)^2, E the batch-mean power of the
    real 2-D transform.  An MSE fit leaves out the part of the forcing its
    inputs do not determine and with it power at every scale; this term asks
    for the target's power scale by scale.  (torch autograd on the output
    only; the network's own backward pass takes the result.)"""
    nch = target.shape[-1]
    o = out[..., :nch].detach().clone().requires_grad_(True)
    with torch.enable_grad():
        po = torch.fft.rfft2(o.permute(0, 3, 1, 2))
        pt = torch.fft.rfft2(target.permute(0, 3, 1, 2))
        eo = (po.real ** 2 + po.imag ** 2).mean(dim=0)
        et = (pt.real ** 2 + pt.imag ** 2).mean(dim=0)
        eps = 1e-6 * et.mean()
        loss = ((torch.log(eo + eps) - torch.log(et + eps)) ** 2).mean()
        g, = torch.autograd.grad(loss, o)
    full = torch.zeros_like(out)
    full[..., :nch] = g
    return full


# ---------------------------------------------------------------- network
class CNN:
    """5x5 periodic convolutions with ReLU between them; channels[0] inputs,
    channels[-1] outputs (128/64/32 hidden channels by default)."""

    def __init__(self, channels=DEFAULT_CHANNELS, seed=0, device="cuda"):
        self.channels = tuple(int(c) for c in channels)
        self.chp = [pad16(c) for c in self.channels]
        self.device = device
        g = torch.Generator().manual_seed(seed)
        self.w32, self.b32, self.wb, self.wt = [], [], [], []
        for l in range(len(self.channels) - 1):
            cin, cout = self.channels[l], self.channels[l + 1]
            cinp, coutp = self.chp[l], self.chp[l + 1]
            bound = 1.0 / math.sqrt(cin * 25)          # PyTorch's conv default init
            w = torch.zeros((coutp, 25, cinp))
            w[:cout, :, :cin] = (torch.rand((cout, 25, cin), generator=g) * 2 - 1) * bound
            b = torch.zeros(coutp)
            b[:cout] = (torch.rand(cout, generator=g) * 2 - 1) * bound
            self.w32.append(w.to(device))
            self.b32.append(b.to(device))
            self.wb.append(self.w32[-1].to(torch.bfloat16))
            self.wt.append(torch.empty((cinp, 25, coutp), dtype=torch.bfloat16, device=device))
        self._adam = None
        self.step_count = 0
        # the first and last layers run special kernels when their shapes allow
        self.first_packed = self.channels[0] <= 8
        self.last_packed = 5 * self.channels[-1] <= 16 and self.chp[-2] % 32 == 0
        self.w_first = self.w_last = None
        self._repack()

    @property
    def nlayers(self):
        return len(self.channels) - 1

    def n_params(self):
        return sum((self.channels[l + 1] * self.channels[l] * 25 + self.channels[l + 1])
                   for l in range(self.nlayers))

    def refresh_bf16(self):
        for l in range(self.nlayers):
            self.wb[l] = self.w32[l].to(torch.bfloat16)
        self._repack()

    def _repack(self):
        """The first and last layers' weights in their special kernels' layouts
        (after every change of the BF16 weights)."""
        if self.first_packed:
            self.w_first = pack_first(self.wb[0], self.w_first)
        if self.last_packed:
            self.w_last = pack_last(self.wb[-1], self.channels[-1], self.w_last)

    def _layer(self, l, h, relu, out_bf16, out_f32, y=None, y32=None, stream=None):
        """Forward of layer l on the kernel that fits its shape."""
        coutp = self.chp[l + 1]
        if l == 0 and self.first_packed:
            return conv_first(h, self.w_first, self.b32[0], coutp, relu=relu, y=y, y32=y32,
                              out_bf16=out_bf16, out_f32=out_f32, stream=stream)
        if l == self.nlayers - 1 and self.last_packed and not relu:
            if y32 is not None:  # a preallocated output may be narrower (out_width)
                coutp = y32.shape[-1]
            return conv_last(h, self.w_last, self.b32[l], coutp, self.channels[-1], y=y, y32=y32,
                             out_bf16=out_bf16, out_f32=out_f32, stream=stream)
        return conv_fwd(h, self.wb[l], self.b32[l], coutp, relu=relu, out_bf16=out_bf16,
                        out_f32=out_f32, y=y, y32=y32, stream=stream)

    def pad_input(self, x):
        """(B, H, W, channels[0]) float -> padded bf16 NHWC."""
        B, H, W, c = x.shape
        xp = torch.zeros((B, H, W, self.chp[0]), dtype=torch.bfloat16, device=x.device)
        xp[..., :c] = x.to(torch.bfloat16)
        return xp

    def forward(self, xp, keep=False):
        """xp: padded bf16 input.  Returns the FP32 output (B,H,W,coutp) and,
        with keep, every BF16 activation for the backward pass."""
        acts = [xp]
        h = xp
        for l in range(self.nlayers):
            last = l == self.nlayers - 1
            y, y32 = self._layer(l, h, relu=not last, out_bf16=not last or keep, out_f32=last)
            if not last:
                h = y
                if keep:
                    acts.append(y)
            else:
                out = y32
                if keep:
                    acts.append(y)
        return (out, acts) if keep else out

    def out_width(self):
        """Channels of forward_into's FP32 output: the last-layer kernel writes
        4 (of which channels[-1] are real), any other kernel the padded count."""
        return 4 if self.last_packed and self.channels[-1] <= 4 else self.chp[-1]

    def activation_buffers(self, B, H, W):
        """Persistent outputs of every layer for forward_into: BF16 for the
        hidden layers, FP32 for the last (out_width channels)."""
        return [torch.empty((B, H, W, self.out_width() if l == self.nlayers - 1
                             else self.chp[l + 1]),
                            dtype=torch.float32 if l == self.nlayers - 1 else torch.bfloat16,
                            device=self.device) for l in range(self.nlayers)]

    def forward_into(self, xp, bufs, stream):
        """The forward pass into preallocated buffers (activation_buffers) on a
        given stream: nothing is allocated, so it can be recorded into a graph."""
        h = xp
        for l in range(self.nlayers):
            last = l == self.nlayers - 1
            self._layer(l, h, relu=not last, out_bf16=not last, out_f32=last,
                        y=None if last else bufs[l], y32=bufs[l] if last else None,
                        stream=stream)
            h = bufs[l]
        return bufs[-1]

    def backward(self, acts, dout):
        """dout: bf16 gradient of the loss w.r.t. the last (pre-activation) output.
        Returns FP32 (dW, db) per layer."""
        # every gradient is a view of one zeroed buffer (segments on 256-byte
        # boundaries), so that the norm and the clip are one launch each
        def seg(n):
            return (n + 63) // 64 * 64
        sizes = [(self.w32[l].numel(), self.b32[l].numel()) for l in range(self.nlayers)]
        flat = torch.zeros(sum(seg(a) + seg(b) for a, b in sizes), dtype=torch.float32,
                           device=dout.device)
        views, o = [], 0
        for l, (a, b) in enumerate(sizes):
            dw_l = flat[o:o + a].view(self.w32[l].shape)
            o += seg(a)
            db_l = flat[o:o + b].view(self.b32[l].shape)
            o += seg(b)
            views.append((dw_l, db_l))
        self._grad_flat = flat
        dws, dbs = [], []
        d = dout
        for l in reversed(range(self.nlayers)):
            dw, db = views[l]
            conv_wgrad(d, acts[l], dw)
            bias_grad(d, db)
            dws.append(dw)
            dbs.append(db)
            if l > 0:
                flip_transpose(self.wb[l], self.wt[l])
                d, _ = conv_fwd(d, self.wt[l], None, self.chp[l], relu=False, mask=acts[l])
        return dws[::-1], dbs[::-1]

    def adam_step(self, dws, dbs, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        if self._adam is None:
            self._adam = [(torch.zeros_like(w), torch.zeros_like(w), torch.zeros_like(b),
                           torch.zeros_like(b)) for w, b in zip(self.w32, self.b32)]
        self.step_count += 1
        for l in range(self.nlayers):
            mw, vw, mb, vb = self._adam[l]
            check(lib.qgcnn_adam(_p(self.w32[l]), _p(dws[l]), _p(mw), _p(vw), _p(self.wb[l]),
                                 self.w32[l].numel(), lr, b1, b2, eps, self.step_count, wd,
                                 _stream()))
            check(lib.qgcnn_adam(_p(self.b32[l]), _p(dbs[l]), _p(mb), _p(vb), None,
                                 self.b32[l].numel(), lr, b1, b2, eps, self.step_count, 0.0,
                                 _stream()))
        self._repack()

    def train_step(self, xp, target, lr=1e-3, clip=None, spec_weight=0.0):
        """One Adam step on the MSE between output channels and target
        (B,H,W,channels[-1]) FP32; returns the loss tensor.  The global
        gradient norm is kept in `last_grad_norm` (a device scalar); with
        `clip` the gradients are scaled down to at most that norm first.
        spec_weight > 0 adds that times spectral_loss to the loss."""
        out, acts = self.forward(xp, keep=True)
        grad = torch.empty(out.shape, dtype=torch.bfloat16, device=out.device)
        loss = mse(out, target, grad)
        if spec_weight:
            g = spectral_loss_grad(out, target)
            grad = (grad.float() + spec_weight * g).to(torch.bfloat16)
        dws, dbs = self.backward(acts, grad)
        # the gradients are views of one buffer whose padding stays zero
        gn = torch.linalg.vector_norm(self._grad_flat)
        self.last_grad_norm = gn
        if clip is not None:
            self._grad_flat.mul_(torch.clamp(clip / (gn + 1e-30), max=1.0))
        self.adam_step(dws, dbs, lr=lr)
        return loss

    # ------------------------------------------------------------ persistence
    def state(self):
        return {"channels": np.array(self.channels),
                **{f"w{l}": self.w32[l].cpu().numpy() for l in range(self.nlayers)},
                **{f"b{l}": self.b32[l].cpu().numpy() for l in range(self.nlayers)}}

    def save(self, path):
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path, device="cuda"):
        d = np.load(path, allow_pickle=False)
        net = cls(tuple(int(c) for c in d["channels"]), device=device)
        for l in range(net.nlayers):
            net.w32[l] = torch.from_numpy(d[f"w{l}"]).to(device)
            net.b32[l] = torch.from_numpy(d[f"b{l}"]).to(device)
        net.refresh_bf16()
        return net
