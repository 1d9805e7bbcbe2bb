"""The trained CNN as a pyqg q parameterization.

Two ways in:
  * `CNNParameterization(net, scales)` is a callable with
    parameterization_type = "q_parameterization" (pyqg/parameterizations.py:
    4-130): m -> dq (nz, ny, nx), usable by pyqg_amd's models and by pyqg's own
    (the host-callback path of kernel.pyx:424-442);
  * `ClosedEnsemble` steps an ensemble of coarse models with the network (or
    one of the built-in closures of pyqg_amd.parameterizations) inside the
    step on the GPU: per step the inversion, the packing of q, u, v
    into the network input, the network, and dqhdt += rfft2(S) run as device
    kernels with no host round trip, in the reference's step order
    (pyqg/model.py:406-436: invert, advection, friction, q parameterization,
    timestep).
"""
import ctypes

import numpy as np
import torch

from .. import _lib
from .. import constants as C
from .cnn import CNN


class Scales:
    """Per-channel normalisation: network input = field / in_scale, forcing =
    network output * out_scale."""

    def __init__(self, in_scale, out_scale):
        self.in_scale = np.asarray(in_scale, dtype=np.float64)
        self.out_scale = np.asarray(out_scale, dtype=np.float64)

    @classmethod
    def fit(cls, x, y):
        """RMS of every input and target channel of training tensors (NHWC)."""
        xs = torch.sqrt((x.double() ** 2).mean(dim=(0, 1, 2))).cpu().numpy()
        ys = torch.sqrt((y.double() ** 2).mean(dim=(0, 1, 2))).cpu().numpy()
        return cls(xs, ys)


class CNNParameterization:
    """q parameterization from a trained network (host callback path)."""

    parameterization_type = "q_parameterization"

    def __init__(self, net: CNN, scales: Scales):
        self.net = net
        self.scales = scales

    def predict(self, x):
        """x: (B, H, W, 6) FP32 physical inputs on the GPU -> (B, H, W, 2) FP32."""
        xin = x / torch.as_tensor(self.scales.in_scale, dtype=torch.float32, device=x.device)
        out = self.net.forward(self.net.pad_input(xin))
        nout = self.net.channels[-1]
        return out[..., :nout] * torch.as_tensor(self.scales.out_scale, dtype=torch.float32,
                                                 device=x.device)

    def __call__(self, m):
        q, u, v = (np.asarray(a) for a in (m.q, m.u, m.v))
        x = np.concatenate([q, u, v], axis=0).transpose(1, 2, 0)[np.newaxis]
        y = self.predict(torch.as_tensor(x, dtype=torch.float32, device="cuda"))
        return y[0].permute(2, 0, 1).double().cpu().numpy()


class CNNAverageParameterization(CNNParameterization):
    """The mean of several trained networks' forcings (same inputs and
    scales): averaging independently trained networks narrows the spread the
    training seed leaves in the learned closure.  `net` is the first network
    (its input layout is the common one); the closed ensemble runs every
    network and adds each forcing with weight 1/len(nets)."""

    def __init__(self, nets, scales: Scales):
        if not nets:
            raise ValueError("no networks to average")
        if any(n.chp[0] != nets[0].chp[0] or n.channels[-1] != nets[0].channels[-1]
               for n in nets):
            raise ValueError("the averaged networks need the same input and output channels")
        super().__init__(nets[0], scales)
        self.nets = list(nets)

    def predict(self, x):
        xin = x / torch.as_tensor(self.scales.in_scale, dtype=torch.float32, device=x.device)
        nout = self.net.channels[-1]
        out = sum(n.forward(n.pad_input(xin))[..., :nout] for n in self.nets) / len(self.nets)
        return out * torch.as_tensor(self.scales.out_scale, dtype=torch.float32, device=x.device)


class ClosedEnsemble:
    """Coarse two-layer QG ensemble stepped with the CNN closure on the GPU.

    With `param=None` it is the plain coarse model (fused steps), the baseline
    the closure is compared with.  With the network, `fused=True` (the default
    where the grid allows it: two layers, n <= 128) steps through the fused
    two-kernel step: the row kernel writes the network input, the column
    kernel adds the network's forcing to the tendency (qgk_fused_row_net,
    qgk_fused_col_forcing); `fused=False` runs the reference's phases one by
    one with the forcing added after the friction (the same arithmetic up to
    the transforms' rounding)."""

    def __init__(self, nmembers, n, dt, phys, param=None, seed=0, fused=True):
        from ..ensemble import Ensemble
        from ..parameterizations import Parameterization
        self.builtin = isinstance(param, Parameterization)
        uv = self.builtin and param.parameterization_type == 'uv_parameterization'
        self.ens = Ensemble.qg(np.full(nmembers, n), dt=dt, full_state=True,
                               q_param=param is not None and not uv, uv_param=uv, **phys)
        self.b = self.ens.buckets[0]
        self.param = param
        self.n = n
        rng = np.random.default_rng(seed)
        self.ens.set_q([1e-7 * rng.standard_normal((2, n, n)) for _ in range(nmembers)])
        self.fused = False
        if param is not None and not self.builtin:
            self.cinp = param.net.chp[0]
            # zeros: the fused row kernel writes only the 3*nz input channels
            self.xin = torch.zeros((nmembers, n, n, self.cinp), dtype=torch.bfloat16,
                                   device="cuda")
            self.nets = getattr(param, "nets", [param.net])
            # several networks: their forcings are added one by one after the
            # friction (the phase-by-phase form)
            self.fused = (bool(fused) and n <= 128 and self.cinp % 8 == 0 and
                          len(self.nets) == 1)
            if self.fused:
                self.b.set_layer_split(False)
                self.fused = self.b.exchange_columns() == n // 2
            self.actss = [net.activation_buffers(nmembers, n, n) for net in self.nets]
            self.acts = self.actss[0]
            self.inv_in = np.ascontiguousarray(1.0 / param.scales.in_scale)
            self.out_scale = np.ascontiguousarray(param.scales.out_scale)
            self.out_scale_k = np.ascontiguousarray(param.scales.out_scale / len(self.nets))
            self.stream = ctypes.c_void_p(_lib.lib.qgk_stream(self.b.h))
            self._graph_key = None

    @property
    def t(self):
        return self.ens.t

    def set_noise(self, amp, phi=0.0, seed=0):
        """Stochastic forcing beside the network's (qgk_set_spectral_noise):
        amp (nz, n, n/2 + 1) the stationary rms per spectral point in dqh
        units, phi the one-step autocorrelation; None switches it off."""
        if self.param is None or self.builtin:
            raise ValueError("the stochastic forcing goes with the network closure")
        if len(self.nets) > 1:
            raise ValueError("the stochastic forcing goes with a single network (each "
                             "network's forcing add would draw it again)")
        self.b.set_spectral_noise(amp, phi, seed)
        self._graph_key = None

    def step(self, nsteps=1):
        if self.param is None:
            self.ens.step(nsteps)
            return
        if self.builtin:
            self.ens.step_parameterized(nsteps, self.param)
            return
        # the network runs on the bucket's stream into persistent buffers, so
        # once the AB weights are constant two steps are recorded as a HIP
        # graph and replayed (the same kernels as the step-by-step form)
        L = _lib.lib
        h = self.b.h
        nsteps = int(nsteps)
        while nsteps > 0 and self.ens.ablevel < 2:
            self._cnn_step()          # AB start-up
            nsteps -= 1
        if nsteps >= 2:
            w, self.ens.ablevel = C.ab_weights(self.ens.ablevel, self.ens.dt)
            key = (tuple(t.data_ptr() for net in self.nets for t in net.wb),
                   tuple(t.data_ptr() for net in self.nets for t in net.b32), tuple(w))
            ok = self._graph_key == key
            if not ok or L.qgk_graph_replay(h, 0) != 0:
                # allocations and uploads before a capture; the history phase
                # of a kept graph
                self._cnn_step()
                nsteps -= 1
                ok = ok and L.qgk_graph_replay(h, 0) == 0
            if not ok and nsteps >= 2:
                self._graph_key = None
                _lib.check(L.qgk_capture_begin(h))
                try:
                    for _ in range(2):
                        self._cnn_phases(w)
                except Exception:
                    L.qgk_capture_end(h, 0)
                    raise
                _lib.check(L.qgk_capture_end(h, 1))
                self._graph_key = key
                ok = True
            if ok and nsteps >= 2:
                reps = nsteps // 2
                _lib.check(L.qgk_graph_replay(h, reps))
                for _ in range(2 * reps):
                    self.ens.t += self.ens.dt
                self.ens.tc += 2 * reps
                nsteps -= 2 * reps
        for _ in range(nsteps):
            self._cnn_step()

    def _cnn_phases(self, w):
        """Invert, advect, friction, the network's forcing, time step (weights w)."""
        L = _lib.lib
        h = self.b.h
        if self.fused:
            _lib.check(L.qgk_fused_row_net(h, _lib.ptr(self.inv_in), self.cinp,
                                           ctypes_p(self.xin)))
            out = self.param.net.forward_into(self.xin, self.acts, self.stream)
            _lib.check(L.qgk_fused_col_forcing(h, ctypes_p(out), out.shape[-1],
                                               _lib.ptr(self.out_scale), *w))
            return
        _lib.check(L.qgk_invert(h))
        _lib.check(L.qgk_advect(h))
        _lib.check(L.qgk_friction(h))
        _lib.check(L.qgk_pack_inputs(h, _lib.ptr(self.inv_in), self.cinp, ctypes_p(self.xin)))
        scale = self.out_scale if len(self.nets) == 1 else self.out_scale_k
        for net, acts in zip(self.nets, self.actss):
            out = net.forward_into(self.xin, acts, self.stream)
            _lib.check(L.qgk_add_forcing_device(h, ctypes_p(out), out.shape[-1],
                                                _lib.ptr(scale)))
        _lib.check(L.qgk_timestep(h, *w))

    def _cnn_step(self):
        w, self.ens.ablevel = C.ab_weights(self.ens.ablevel, self.ens.dt)
        self._cnn_phases(w)
        self.ens.t += self.ens.dt
        self.ens.tc += 1

    def ph(self):
        return self.b.get("ph")


def ctypes_p(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())
