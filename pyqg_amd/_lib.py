"""ctypes binding of libqgk (include/pyqg_amd.h).

The shared library is built in-tree by pyqg_amd.build.  There is no fallback:
if it is missing or cannot be loaded, importing this module raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_qgk.so")

# error codes
E_ARG, E_HIP, E_STATE, E_NOMEM = -1, -2, -3, -4
# flags
FULL_STATE, HAS_Q_PARAM, HAS_UV_PARAM = 1, 2, 4
# field ids
FIELDS = ["q", "qh", "ph", "u", "v", "uh", "vh", "uq", "vq", "uqh", "vqh", "dqhdt",
          "dqhdt_p", "dqhdt_pp", "dq", "dqh", "du", "dv", "duh", "dvh"]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
REAL_FIELDS = {"q", "u", "v", "uq", "vq", "dq", "du", "dv"}
# constant ids
CONSTS = ["kk", "ll", "k2l2", "filtr", "a", "Ubg", "Qy"]
CONST_ID = {name: i for i, name in enumerate(CONSTS)}

EXPORTS = ["qgk_create", "qgk_destroy", "qgk_last_error", "qgk_set_const", "qgk_set_const_all",
           "qgk_set_inversion_2layer", "qgk_set_inversion_tridiagonal",
           "qgk_get_inversion_tridiagonal", "qgk_set_rek", "qgk_put", "qgk_get", "qgk_invert",
           "qgk_advect", "qgk_friction", "qgk_add_q_param", "qgk_add_uv_param",
           "qgk_timestep", "qgk_q_to_qh", "qgk_qh_to_q", "qgk_fft", "qgk_ifft", "qgk_step",
           "qgk_timing_enable", "qgk_timing", "qgk_sync", "qgk_step_buckets",
           "qgk_copy_probe", "qgk_pattern_probe", "qgk_exchange_columns", "qgk_set_packed_exchange", "qgk_capture_begin", "qgk_capture_end",
           "qgk_graph_replay", "qgk_stream", "qgk_set_variant", "qgk_get_variant", "qgk_set_layer_split",
           "qgk_get_layer_split", "qgk_autotune",
           "qgk_device_name", "qgk_put_device", "qgk_get_device", "qgk_coarsen",
           "qgk_subgrid_forcing", "qgk_pack_inputs", "qgk_add_forcing_device",
           "qgk_fused_row_net", "qgk_fused_col_forcing", "qgk_set_spectral_noise",
           "qgk_diag_set_params", "qgk_diag_accumulate", "qgk_diag_read", "qgk_diag_reset",
           "qgk_param_smagorinsky", "qgk_smagorinsky_viscosity", "qgk_param_zb2020",
           "qgk_param_backscatter", "qgk_apply_q_param", "qgk_apply_uv_param",
           "qgk_member_speed", "qgk_mask_member", "qgk_member_speed_noalloc"]

# device diagnostics, in the order of the qgk_diag enum
DIAGS = ["Ensspec", "KEspec", "EKEdiss", "KEfrictionspec", "EKE", "Dissspec", "ENSDissspec",
         "paramspec", "ENSparamspec", "entspec", "paramspec_APEflux", "paramspec_KEflux",
         "ENSflux", "ENSgenspec", "ENSfrictionspec", "APEgenspec", "APEflux", "KEflux", "APEgen",
         "KEspec_modal", "PEspec_modal", "APEspec", "KEflux_div", "APEflux_div"]
DIAG_ID = {name: i for i, name in enumerate(DIAGS)}
SCALAR_DIAGS = {"EKEdiss", "EKE", "APEgen"}


class QGKError(RuntimeError):
    pass


def _one_hip_runtime():
    """Keep a single HIP runtime in the process.

    The PyTorch-ROCm wheel ships its own libamdhip64.so; libqgk needs
    libamdhip64.so.7 from the ROCm install.  If both copies are mapped, each
    brings its own HSA runtime and whichever initialises second finds no
    device.  When torch is already imported its copy is loaded and satisfies
    libqgk's dependency (soname libamdhip64.so.7).  Otherwise the ROCm copy is
    loaded here under the name torch links against (libamdhip64.so), so a
    later `import torch` binds to it instead of loading its own."""
    import sys
    if "torch" in sys.modules:
        return
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            ctypes.CDLL(name, mode=ctypes.RTLD_GLOBAL)
            return
        except OSError:
            continue


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libqgk is not built ({LIB_PATH} missing): run `python pyqg_amd/build.py` "
            "(needs hipcc); pyqg_amd has no CPU fallback")
    _one_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D, L = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    DP = ctypes.POINTER(ctypes.c_double)
    sig = {
        "qgk_create": (I, [I, I, I, I, I, ctypes.POINTER(P)]),
        "qgk_destroy": (I, [P]),
        "qgk_last_error": (ctypes.c_char_p, []),
        "qgk_set_const": (I, [P, I, I, P]),
        "qgk_set_const_all": (I, [P, I, P]),
        "qgk_set_inversion_2layer": (I, [P, I, D, D]),
        "qgk_set_inversion_tridiagonal": (I, [P, I, DP]),
        "qgk_get_inversion_tridiagonal": (I, [P, I, ctypes.POINTER(I)]),
        "qgk_set_rek": (I, [P, I, D]),
        "qgk_put": (I, [P, I, P]),
        "qgk_get": (I, [P, I, P]),
        "qgk_invert": (I, [P]),
        "qgk_advect": (I, [P]),
        "qgk_friction": (I, [P]),
        "qgk_add_q_param": (I, [P, P]),
        "qgk_add_uv_param": (I, [P, P, P]),
        "qgk_timestep": (I, [P, D, D, D]),
        "qgk_q_to_qh": (I, [P]),
        "qgk_qh_to_q": (I, [P]),
        "qgk_fft": (I, [P, I, P, P]),
        "qgk_ifft": (I, [P, I, P, P]),
        "qgk_step": (I, [P, I, P, I]),
        "qgk_timing_enable": (I, [P, I]),
        "qgk_timing": (I, [P, DP, ctypes.POINTER(L), DP, ctypes.POINTER(L)]),
        "qgk_sync": (I, [P]),
        "qgk_step_buckets": (I, [ctypes.POINTER(P), I, I, P, I]),
        "qgk_copy_probe": (I, [ctypes.c_size_t, I, DP]),
        "qgk_pattern_probe": (I, [I, I, I, I, I, I, DP, DP]),
        "qgk_exchange_columns": (I, [P, ctypes.POINTER(I)]),
        "qgk_set_packed_exchange": (I, [P, I]),
        "qgk_capture_begin": (I, [P]),
        "qgk_capture_end": (I, [P, I]),
        "qgk_graph_replay": (I, [P, I]),
        "qgk_stream": (P, [P]),
        "qgk_set_variant": (I, [P, I, I]),
        "qgk_get_variant": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "qgk_set_layer_split": (I, [P, I]),
        "qgk_get_layer_split": (I, [P, ctypes.POINTER(I)]),
        "qgk_autotune": (I, [P, I, ctypes.POINTER(I), ctypes.POINTER(I), DP, DP]),
        "qgk_device_name": (I, [ctypes.c_char_p, I]),
        "qgk_put_device": (I, [P, I, P]),
        "qgk_get_device": (I, [P, I, P]),
        "qgk_coarsen": (I, [P, P, I, I, P]),
        "qgk_subgrid_forcing": (I, [P, P, P]),
        "qgk_pack_inputs": (I, [P, P, I, P]),
        "qgk_fused_row_net": (I, [P, P, I, P]),
        "qgk_fused_col_forcing": (I, [P, P, I, P, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double]),
        "qgk_add_forcing_device": (I, [P, P, I, P]),
        "qgk_set_spectral_noise": (I, [P, P, ctypes.c_double, ctypes.c_ulonglong]),
        "qgk_diag_set_params": (I, [P, I, P]),
        "qgk_diag_accumulate": (I, [P, ctypes.c_uint, I, D, D, D, D]),
        "qgk_diag_read": (I, [P, I, I, I, P]),
        "qgk_diag_reset": (I, [P]),
        "qgk_param_smagorinsky": (I, [P, P, D, D, I, P, P]),
        "qgk_smagorinsky_viscosity": (I, [P, P, D, P]),
        "qgk_param_zb2020": (I, [P, D, D, I, P, P]),
        "qgk_param_backscatter": (I, [P, P, P, D, D, D, D, I, P]),
        "qgk_apply_q_param": (I, [P]),
        "qgk_apply_uv_param": (I, [P]),
        "qgk_member_speed": (I, [P, P]),
        "qgk_mask_member": (I, [P, I]),
        "qgk_member_speed_noalloc": (I, [P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def check(rc):
    if rc == 0:
        return
    msg = lib.qgk_last_error().decode(errors="replace")
    if rc == E_ARG:
        raise ValueError(msg)
    if rc == E_STATE:
        raise AttributeError(msg)
    if rc == E_NOMEM:
        raise MemoryError(msg)
    raise QGKError(msg)


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Bucket:
    """Owning wrapper of one qgk_bucket handle."""

    def __init__(self, nmembers, nz, ny, nx, flags):
        h = ctypes.c_void_p()
        check(lib.qgk_create(int(nmembers), int(nz), int(ny), int(nx), int(flags),
                             ctypes.byref(h)))
        self.h = h
        self.flags = int(flags)
        self.M, self.nz, self.ny, self.nx = int(nmembers), int(nz), int(ny), int(nx)
        self.nl, self.nk = self.ny, self.nx // 2 + 1

    def close(self):
        if getattr(self, "h", None):
            lib.qgk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, field):
        if field in REAL_FIELDS:
            return (self.M, self.nz, self.ny, self.nx)
        return (self.M, self.nz, self.nl, self.nk)

    def put(self, field, arr):
        dtype = np.float64 if field in REAL_FIELDS else np.complex128
        a = np.ascontiguousarray(arr, dtype=dtype)
        if a.size != int(np.prod(self.shape(field))):
            raise ValueError(f"{field}: expected {self.shape(field)}, got {np.shape(arr)}")
        check(lib.qgk_put(self.h, FIELD_ID[field], ptr(a)))

    def get(self, field, out=None):
        dtype = np.float64 if field in REAL_FIELDS else np.complex128
        if out is None:
            out = np.empty(self.shape(field), dtype)
        assert out.flags.c_contiguous and out.dtype == dtype
        check(lib.qgk_get(self.h, FIELD_ID[field], ptr(out)))
        return out

    def set_const(self, name, arr, member=-1):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        check(lib.qgk_set_const(self.h, CONST_ID[name], int(member), ptr(a)))

    def set_const_all(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        check(lib.qgk_set_const_all(self.h, CONST_ID[name], ptr(a)))

    def set_inversion_2layer(self, F1, F2, member=-1):
        check(lib.qgk_set_inversion_2layer(self.h, int(member), float(F1), float(F2)))

    def set_inversion_tridiagonal(self, S, member=-1):
        """Fused steps solve (S - kappa^2 I) ph = qh with the tridiagonal
        stretching matrix S (nz x nz) instead of reading the table a; None
        switches the solve off."""
        if S is None:
            check(lib.qgk_set_inversion_tridiagonal(self.h, int(member), None))
            return
        S = np.asarray(S, dtype=np.float64)
        nz = S.shape[0]
        if S.shape != (nz, nz) or np.any(np.triu(S, 2)) or np.any(np.tril(S, -2)):
            raise ValueError("S must be a square tridiagonal matrix")
        bands = np.zeros((3, nz))
        bands[0, 1:] = np.diag(S, -1)
        bands[1] = np.diag(S)
        bands[2, :-1] = np.diag(S, 1)
        bands = np.ascontiguousarray(bands)
        check(lib.qgk_set_inversion_tridiagonal(
            self.h, int(member), bands.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))

    def inversion_tridiagonal(self, member=0):
        on = ctypes.c_int()
        check(lib.qgk_get_inversion_tridiagonal(self.h, int(member), ctypes.byref(on)))
        return bool(on.value)

    def set_rek(self, rek, member=-1):
        check(lib.qgk_set_rek(self.h, int(member), float(rek)))

    def step(self, coeffs, use_graph=False):
        c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, 3)
        check(lib.qgk_step(self.h, c.shape[0], ptr(c), int(bool(use_graph))))

    def timing_enable(self, on=True):
        check(lib.qgk_timing_enable(self.h, int(bool(on))))

    def timing(self):
        mr, mc = ctypes.c_double(), ctypes.c_double()
        nr, nc = ctypes.c_long(), ctypes.c_long()
        check(lib.qgk_timing(self.h, ctypes.byref(mr), ctypes.byref(nr), ctypes.byref(mc),
                             ctypes.byref(nc)))
        return mr.value, nr.value, mc.value, nc.value

    def sync(self):
        check(lib.qgk_sync(self.h))

    def member_speed(self):
        """Largest |u + Ubg|, |v| over each member's grid (NaN or inf when the
        member's state is not finite)."""
        out = np.empty(self.M)
        check(lib.qgk_member_speed(self.h, ptr(out)))
        return out

    def mask_member(self, member):
        check(lib.qgk_mask_member(self.h, int(member)))

    def set_variant(self, erow, ecol):
        check(lib.qgk_set_variant(self.h, int(erow), int(ecol)))

    def set_layer_split(self, on):
        check(lib.qgk_set_layer_split(self.h, int(bool(on))))

    def layer_split(self):
        on = ctypes.c_int()
        check(lib.qgk_get_layer_split(self.h, ctypes.byref(on)))
        return bool(on.value)

    def variant(self):
        r, c = ctypes.c_int(), ctypes.c_int()
        check(lib.qgk_get_variant(self.h, ctypes.byref(r), ctypes.byref(c)))
        return r.value, c.value

    def set_packed_exchange(self, on):
        check(lib.qgk_set_packed_exchange(self.h, int(bool(on))))

    def set_spectral_noise(self, amp, phi=0.0, seed=0):
        """Stochastic forcing added to the network's each step (AR(1) per
        spectral point, amp (nz, ny, nk) its stationary rms in dqh units, phi
        the one-step autocorrelation); amp None switches it off."""
        if amp is None:
            check(lib.qgk_set_spectral_noise(self.h, None, 0.0, 0))
            return
        a = np.ascontiguousarray(amp, dtype=np.float64)
        if a.shape != (self.nz, self.ny, self.nk):
            raise ValueError(f"amp must have shape {(self.nz, self.ny, self.nk)}")
        check(lib.qgk_set_spectral_noise(self.h, a.ctypes.data_as(ctypes.c_void_p), float(phi),
                                         int(seed)))

    def exchange_columns(self):
        """Columns per row of the x-mixed fields the fused kernels exchange:
        nx/2 on the packed register path, nx/2 + 1 otherwise."""
        n = ctypes.c_int()
        check(lib.qgk_exchange_columns(self.h, ctypes.byref(n)))
        return n.value

    # ---------------------------------------------------------- diagnostics
    def diag_slots(self, name):
        """Leading shape of one member's value of a device diagnostic."""
        nz = self.nz
        if name in ("Ensspec", "KEspec", "KEspec_modal", "EKE"):
            return nz
        if name == "PEspec_modal":
            return nz - 1
        return 1

    def diag_set_params(self, member, params):
        a = np.ascontiguousarray(params, dtype=np.float64)
        n = 2 * self.nz * self.nz + 5 * self.nz + 5
        if a.size != n:
            raise ValueError(f"diagnostic parameters: expected {n} values, got {a.size}")
        check(lib.qgk_diag_set_params(self.h, int(member), ptr(a)))

    def diag_accumulate(self, names, dt, weights, instant=False):
        mask = 0
        for n in names:
            mask |= 1 << DIAG_ID[n]
        check(lib.qgk_diag_accumulate(self.h, mask, int(bool(instant)), float(dt),
                                      *(float(w) for w in weights)))

    def diag_read(self, name, member=-2, instant=False):
        """Accumulated sum of one diagnostic: member >= 0 one member, -1 all
        members (leading axis M), -2 the sum over members.  Shapes follow the
        reference: (nz, nl, nk), (nl, nk), (nz,) or a scalar."""
        slots = self.diag_slots(name)
        scalar = name in SCALAR_DIAGS
        one = (slots,) if scalar else (slots, self.nl, self.nk)
        shape = (self.M,) + one if member == -1 else one
        out = np.empty(shape, np.float64)
        check(lib.qgk_diag_read(self.h, DIAG_ID[name], int(member), int(bool(instant)),
                                ptr(out)))
        squeeze = (name not in ("Ensspec", "KEspec", "KEspec_modal", "EKE", "PEspec_modal"))
        if squeeze:
            out = out.reshape(out.shape[:-3] + out.shape[-2:]) if not scalar else \
                out.reshape(out.shape[:-1])
        return out

    def diag_reset(self):
        check(lib.qgk_diag_reset(self.h))

    def autotune(self, reps=5):
        r, c = ctypes.c_int(), ctypes.c_int()
        mr, mc = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        check(lib.qgk_autotune(self.h, int(reps), ctypes.byref(r), ctypes.byref(c), mr, mc))
        return (r.value, c.value), list(mr), list(mc)


def device_name():
    buf = ctypes.create_string_buffer(256)
    check(lib.qgk_device_name(buf, 256))
    return buf.value.decode()


def copy_probe(nbytes, iters=20):
    """Milliseconds per launch of a plain copy kernel moving nbytes in and out."""
    ms = ctypes.c_double()
    check(lib.qgk_copy_probe(int(nbytes), int(iters), ctypes.byref(ms)))
    return ms.value


def pattern_probe(ny, ldk, members, nz, iters=10, ncols=None):
    """(ms_rows, ms_cols): memory-only stand-ins for the fused step's row and
    column kernels at a bucket's shape (same fields, bytes and access
    patterns, no transforms); ncols = columns touched per row (default ldk)."""
    r, c = ctypes.c_double(), ctypes.c_double()
    check(lib.qgk_pattern_probe(int(ny), int(ldk), int(ncols or ldk), int(members), int(nz),
                                int(iters), ctypes.byref(r), ctypes.byref(c)))
    return r.value, c.value


def step_buckets(buckets, coeffs, use_graph=False):
    c = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, 3)
    arr = (ctypes.c_void_p * len(buckets))(*[b.h for b in buckets])
    check(lib.qgk_step_buckets(arr, len(buckets), c.shape[0], ptr(c), int(bool(use_graph))))
