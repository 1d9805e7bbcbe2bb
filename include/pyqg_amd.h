/* C-ABI of the MI355X pyqg time-stepping core (libqgk).
 *
 * The library replaces the hot path of pyqg's `cdef class PseudoSpectralKernel`
 * (pyqg/kernel.pyx:32-637): it owns the model state in HBM and runs the
 * inversion, advection, friction, parameterization adds, AB3 step and the 2-D
 * real FFTs as HIP kernels for gfx950, in FP64/complex128.
 *
 * Handles are "buckets": M independent members ("ensemble members") of one
 * grid size (nz layers, ny x nx points).  A single pyqg model is a bucket with
 * M = 1 created with QGK_FULL_STATE, which also keeps every intermediate field
 * the reference exposes (ph, u, v, uh, vh, uq, vq, uqh, vqh, ...).
 *
 * Layouts (C order, host and device alike):
 *   real field     double [M][nz][ny][nx]
 *   spectral field double complex (interleaved re,im) [M][nz][nl][nk],
 *                  nl = ny, nk = nx/2 + 1
 * nx and ny are 2^a in [8, 1024] or 3 * 2^a in [24, 768].
 *
 * Every function returns 0 on success or a negative QGK_E* code; the message
 * of the last failure on the calling thread is qgk_last_error().  Calls are
 * synchronous with respect to the host unless the name ends in _async.
 */
#ifndef PYQG_AMD_H
#define PYQG_AMD_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qgk_bucket qgk_bucket;

enum {
  QGK_OK = 0,
  QGK_E_ARG = -1,     /* bad argument (maps to ValueError) */
  QGK_E_HIP = -2,     /* HIP runtime failure */
  QGK_E_STATE = -3,   /* field not allocated for this handle (maps to AttributeError) */
  QGK_E_NOMEM = -4
};

/* creation flags */
enum {
  QGK_FULL_STATE = 1, /* keep every reference-visible intermediate field */
  QGK_HAS_Q_PARAM = 2, /* allocate dq, dqh   (kernel.pyx:174-178) */
  QGK_HAS_UV_PARAM = 4 /* allocate du, dv, duh, dvh (kernel.pyx:164-172) */
};

/* field ids for qgk_put / qgk_get (the state arrays of kernel.pyx:37-67) */
enum {
  QGK_Q = 0, QGK_QH, QGK_PH, QGK_U, QGK_V, QGK_UH, QGK_VH, QGK_UQ, QGK_VQ, QGK_UQH,
  QGK_VQH, QGK_DQHDT, QGK_DQHDT_P, QGK_DQHDT_PP, QGK_DQ, QGK_DQH, QGK_DU, QGK_DV,
  QGK_DUH, QGK_DVH, QGK_NFIELDS
};

/* constant ids for qgk_set_const (kernel.pyx:84-104, property setters 528-575) */
enum {
  QGK_KK = 0,   /* double[nk]            zonal wavenumbers */
  QGK_LL,       /* double[nl]            meridional wavenumbers */
  QGK_K2L2,     /* double[nl][nk]        kk^2 + ll^2 (public _k2l2) */
  QGK_FILTR,    /* double[nl][nk]        spectral filter */
  QGK_A,        /* double[nz][nz][nl][nk] real inversion matrix */
  QGK_UBG,      /* double[nz]            background zonal flow */
  QGK_QY,       /* double[nz]            background PV gradient */
  QGK_NCONSTS
};

/* Replaces PseudoSpectralKernel.__init__ (kernel.pyx:120-206) for nmembers
 * members; all fields start at zero, filtr at one, rek at zero. */
int qgk_create(int nmembers, int nz, int ny, int nx, int flags, qgk_bucket** out);
int qgk_destroy(qgk_bucket* b);
const char* qgk_last_error(void);

/* Constant upload.  member < 0: one table shared by every member; member >= 0:
 * that member's own table.  Replaces the kk/ll/_k2l2/filtr/a/Ubg/Qy setters
 * (kernel.pyx:536-575). */
int qgk_set_const(qgk_bucket* b, int which, int member, const double* host);
/* Per-member tables packed member-major: host holds nmembers consecutive
 * tables of constant `which`; member m uses the m-th. */
int qgk_set_const_all(qgk_bucket* b, int which, const double* host);
/* Two-layer closed-form inversion (qg_model.py:146-161) evaluated in the kernel
 * from F1, F2 instead of a table; used by ensembles of QGModels. */
int qgk_set_inversion_2layer(qgk_bucket* b, int member, double F1, double F2);
/* N-layer inversion by a tridiagonal solve (no counterpart in the reference,
 * which multiplies by the table a = (S - kappa^2 I)^-1, layered_model.py:211-229).
 * bands holds the [3][nz] bands of the stretching matrix S (below, on and above
 * the diagonal; bands[0][0] = bands[2][nz-1] = 0), which must be tridiagonal with
 * a non-positive diagonal, non-negative off-diagonals and diagonal dominance, as
 * the layered model's S is.  The fused steps of buckets with 3 <= nz <= 16 then
 * solve (S - kappa^2 I) ph = qh per wavenumber instead of reading a, which stays
 * in use for the per-method phases and must be the inverse of the same S.
 * Writing a (qgk_set_const / qgk_set_const_all) switches the solve off for the
 * members written; bands == NULL switches it off.  member < 0: all members. */
int qgk_set_inversion_tridiagonal(qgk_bucket* b, int member, const double* bands);
/* *on = 1 when member uses the tridiagonal solve (no counterpart in the reference). */
int qgk_get_inversion_tridiagonal(qgk_bucket* b, int member, int* on);
/* Bottom drag rek (public attribute, kernel.pyx:101). member < 0: all. */
int qgk_set_rek(qgk_bucket* b, int member, double rek);

/* Field transfer, all members at once (host array of the full bucket size).
 * Replaces the property getters / setters' copies (kernel.pyx:576-637). */
int qgk_put(qgk_bucket* b, int field, const void* host);
int qgk_get(qgk_bucket* b, int field, void* host);

/* Per-phase operations, one per reference method. */
int qgk_invert(qgk_bucket* b);                  /* kernel.pyx:312-355 */
int qgk_advect(qgk_bucket* b);                  /* kernel.pyx:357-395 */
int qgk_friction(qgk_bucket* b);                /* kernel.pyx:444-461 */
int qgk_add_q_param(qgk_bucket* b, const double* dq_host);                      /* 424-442 */
int qgk_add_uv_param(qgk_bucket* b, const double* du_host, const double* dv_host); /* 397-422 */
/* AB update + filter + history rotation + q = irfft2(qh) (kernel.pyx:463-526).
 * dt1..dt3 are the Adams-Bashforth weights for this step (kernel.pyx:482-500). */
int qgk_timestep(qgk_bucket* b, double dt1, double dt2, double dt3);
/* Step graphs (no counterpart in the reference, whose steps are host loops):
 * between qgk_capture_begin and qgk_capture_end the phase entry points above
 * (qgk_invert, qgk_advect, qgk_friction, qgk_apply_*_param, qgk_timestep) and
 * the built-in closures (qgk_param_*, device targets only) are recorded into
 * a HIP graph instead of run; qgk_graph_replay(b, n) then runs the recorded
 * steps n times with one synchronization.  A graph must hold a multiple of 2
 * steps (the period of the tendency history's buffer swap) and is refused
 * after fused steps (which swap the qh buffers) or a constants
 * change; the closures' dx/Hi must not change while capturing.
 * qgk_capture_end(b, 0) discards the capture. */
int qgk_capture_begin(qgk_bucket* b);
int qgk_capture_end(qgk_bucket* b, int keep);
int qgk_graph_replay(qgk_bucket* b, int n);
/* The handle's HIP stream (hipStream_t), for callers that enqueue their own
 * kernels between the phases (e.g. a network's forward pass inside a captured
 * step).  No counterpart in the reference. */
void* qgk_stream(qgk_bucket* b);
/* The q setter (rfft2 of q into qh, kernel.pyx:576-582) and the qh setter's
 * inverse transform (kernel.pyx:583-591). */
int qgk_q_to_qh(qgk_bucket* b);
int qgk_qh_to_q(qgk_bucket* b);

/* Generic transforms of one member-sized array (nz planes), copy in / copy
 * out (kernel.pyx:291-309).  nplanes lets callers batch any number of planes. */
int qgk_fft(qgk_bucket* b, int nplanes, const double* in_real, double* out_complex);
int qgk_ifft(qgk_bucket* b, int nplanes, const double* in_complex, double* out_real);

/* Fused stepping: nsteps full steps of _step_forward (model.py:406-436) without
 * parameterizations, two kernels per step.  coeffs holds nsteps*3 AB weights
 * (same for every member).  Afterwards qh, q (transformed on first read) and
 * the history are current; the other intermediate fields of QGK_FULL_STATE
 * handles are not (the Python model runs its last step phase by phase so that
 * they are).  use_graph != 0 replays the two-kernel step as a captured
 * hipGraph once the weights stop changing. */
int qgk_step(qgk_bucket* b, int nsteps, const double* coeffs, int use_graph);

/* Per-member health check, the counterpart of the reference's CFL assertion
 * every twrite steps (pyqg/model.py:564-574, pyqg/qg_diagnostics.py:76-79)
 * for every member of a bucket at once: speed_host[m] = max over the grid of
 * |u + Ubg| and |v| of member m's current state (CFL = speed * dt / dx); NaN
 * (or inf) when any of the member's q, u or v is not finite.
 * qgk_mask_member zeroes every state field of one member (its PV, tendency
 * history and intermediates), so a member that blew up rests at q = 0 instead
 * of carrying NaNs into bucket-wide sums. */
int qgk_member_speed(qgk_bucket* b, double* speed_host);
int qgk_mask_member(qgk_bucket* b, int member);
/* qgk_member_speed without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  speed_dev is a
 * caller-provided DEVICE buffer of M doubles; the result equals
 * qgk_member_speed's.  The per-member words live in a static device array of
 * 4096 entries (32 KB per GPU, held until the process ends), so buckets of more
 * than 4096 members are rejected with QGK_E_ARG.  Every call on one GPU shares
 * that array: calls that overlap on one GPU (different streams or threads)
 * race.  Returns after the work is done, except under graph capture of the
 * bucket's stream, where it only enqueues. */
int qgk_member_speed_noalloc(qgk_bucket* b, double* speed_dev);

/* Kernel timing for the measurement harness: when enabled, the fused step
 * records HIP events around each of its kernels; qgk_timing returns the
 * summed milliseconds and launch counts of the row-phase and column-phase
 * kernels since the last reset. */
int qgk_timing_enable(qgk_bucket* b, int enable);
int qgk_timing(qgk_bucket* b, double* ms_row, long* n_row, double* ms_col, long* n_col);
int qgk_sync(qgk_bucket* b);

/* Device-to-device field transfer (dptr is device memory of the full bucket
 * size), for pipelines that keep data on the GPU. */
int qgk_put_device(qgk_bucket* b, int field, const void* dptr);
int qgk_get_device(qgk_bucket* b, int field, void* dptr);

/* Learned-closure support; no counterpart in the reference (pyqg removed its
 * subgrid_forcing, pyqg/model.py:397-402).
 * qgk_coarsen: spectral coarse-graining of a field of `hi` into a field of `lo`
 *   (same members and layers, square grids): the modes both grids resolve,
 *   scaled by (n_lo/n_hi)^2 and multiplied by filter_lo[nl_lo][nk_lo] (NULL:
 *   sharp truncation); the coarse Nyquist row and column are zero.
 * qgk_subgrid_forcing: for the current state of `hi` (full state): the coarse
 *   state C(qh) into lo's qh (and q, u, v after lo's inversion), and the
 *   subgrid PV forcing S = C(T_hi(q)) - T_lo(C(q)) into lo's dqh (spectral) and
 *   dq (real), T = the model tendency of _invert + _do_advection.
 * qgk_pack_inputs: lo's q, u, v (real) as an NHWC BF16 network input with
 *   channels f*nz + z (f = q, u, v) times inv_scale[3*nz], zero-padded to cinp
 *   (a multiple of 8; dst 16-byte aligned).
 * qgk_add_forcing_device: dq = net_out[..., z] * scale[z] (network output FP32
 *   NHWC with coutp channels), then dqhdt += rfft2(dq), as
 *   _do_q_subgrid_parameterization does with a host array (kernel.pyx:424-442). */
int qgk_coarsen(qgk_bucket* hi, qgk_bucket* lo, int field_hi, int field_lo,
                const double* filter_lo);
int qgk_subgrid_forcing(qgk_bucket* hi, qgk_bucket* lo, const double* filter_lo);
int qgk_pack_inputs(qgk_bucket* b, const double* inv_scale, int cinp, void* dst);
int qgk_add_forcing_device(qgk_bucket* b, const float* net_out, int coutp, const double* scale);

/* Fused closed step (no counterpart in the reference: Model._step_forward with
 * a q parameterization, model.py:406-436, as two fused kernels around the
 * network).  One step is qgk_fused_row_net, the network on the bucket's
 * stream (reading net_in, writing net_out), then qgk_fused_col_forcing:
 *   qgk_fused_row_net: the fused row kernel, which also writes the network
 *     input as qgk_pack_inputs does (channels f*nz + z of q, u, v times
 *     inv_scale, NHWC BF16; channels from 3*nz up to cinp are not written and
 *     must hold zeros);
 *   qgk_fused_col_forcing: dq = net_out * scale as qgk_add_forcing_device,
 *     dqh = rfft2(dq), and the fused column kernel with dqh added to the
 *     tendency after the drag, AB weights dt1..dt3; ends the step.
 * Two-layer square grids up to 128 on the packed register kernels (layer split
 * off, standard wavenumber grid); capturable (qgk_capture_begin). */
int qgk_fused_row_net(qgk_bucket* b, const double* inv_scale, int cinp, void* net_in);
int qgk_fused_col_forcing(qgk_bucket* b, const float* net_out, int coutp, const double* scale,
                          double dt1, double dt2, double dt3);

/* Stochastic forcing added to the network's (no counterpart in the reference;
 * for the learned closure): each step, after dqh = rfft2(dq) in
 * qgk_add_forcing_device and qgk_fused_col_forcing,
 *   xi <- phi xi + sqrt(1 - phi^2) amp w,   dqh += xi,
 * per member, layer and spectral point, w complex normal (E|w|^2 = 1) drawn
 * from a counter-based hash of (seed, step, point); the kx = 0 and kx = nx/2
 * columns are drawn Hermitian in l, so xi is the spectrum of a real field when
 * amp is symmetric there.  amp: nz * ny * nk doubles (layer, row, column; the
 * stationary rms of xi, in the units of dqh, i.e. of rfft2 of dq), phi in
 * [0, 1) the one-step autocorrelation.  amp NULL switches it off.  The process
 * restarts from zero at every call; captured steps are recaptured. */
int qgk_set_spectral_noise(qgk_bucket* b, const double* amp, double phi,
                           unsigned long long seed);

/* Two-layer column steps by the layer-split kernel (each layer its own
 * thread group) instead of the register kernel (no counterpart in the
 * reference).  Default: on for one member or members x max(nx, ny)^2 <= 2^19, off
 * otherwise; results agree to rounding.  Error for on with nz != 2. */
int qgk_set_layer_split(qgk_bucket* b, int on);
int qgk_get_layer_split(qgk_bucket* b, int* on);
/* Register radix (4, 8 or 16 complex values per thread) of the fused row and
 * column kernels; it fixes the Stockham pass structure and the tile shapes. */
int qgk_set_variant(qgk_bucket* b, int erow, int ecol);
int qgk_get_variant(qgk_bucket* b, int* erow, int* ecol);
/* Columns per row of the x-mixed fields the fused row and column kernels
 * exchange (no counterpart in the reference): nx/2 when the register column
 * kernels step the bucket (nz <= 2, layer split off; the real x bins 0 and
 * nx/2 share one complex slot; it also needs every member's ll and k2l2 to
 * be the standard tables, which the kernels then form themselves), nx/2 + 1
 * otherwise. */
int qgk_exchange_columns(qgk_bucket* b, int* ncols);
/* Packed exchange on (default) or off for this bucket (no counterpart in the
 * reference): off, the register column kernels exchange nx/2 + 1 columns per
 * row as the other column kernels do.  Results agree to rounding. */
int qgk_set_packed_exchange(qgk_bucket* b, int on);
/* Autotuner: times every radix for the row and the column kernel on this
 * bucket (reps launches each; the model state is saved and restored), selects
 * the fastest, and reports the per-candidate milliseconds (radix 4, 8, 16; -1
 * where a radix exceeds the length).  Not in the reference (no counterpart). */
int qgk_autotune(qgk_bucket* b, int reps, int* erow, int* ecol, double* ms_row3,
                 double* ms_col3);
/* "<name>:<gcnArch>:<CUs>" of the current device, the autotuner cache key. */
int qgk_device_name(char* buf, int len);

/* Measurement aid: average time of a plain device-to-device copy kernel
 * (nbytes read + nbytes written) over iters launches. */
int qgk_copy_probe(size_t nbytes, int iters, double* ms_per_copy);
/* Measurement aid: memory-only stand-ins for the fused step's two kernels at
 * the shape of a bucket (members x nz planes of ny rows with row stride ldk
 * complex128, of which the first ncols are touched, a multiple of 8):
 * ms_rows times three fields read and two written as sequential streams (the
 * row kernel's traffic), ms_cols five fields read and five written by
 * (8-column tile, member) blocks over every row and layer (the column
 * kernel's traffic and access pattern).  Average ms over iters launches;
 * allocates and frees 9 fields. */
int qgk_pattern_probe(int ny, int ldk, int ncols, int members, int nz, int iters,
                      double* ms_rows, double* ms_cols);

/* Multi-bucket stepping for uneven ensembles: nsteps fused steps of every
 * bucket, all launched on the first bucket's stream (one graph when
 * use_graph != 0). */
int qgk_step_buckets(qgk_bucket** bs, int nbuckets, int nsteps, const double* coeffs,
                     int use_graph);

/* Time-averaged diagnostics accumulated on the device, per member, replacing
 * the registry of pyqg/model.py:644-736 (Ensspec ... ENSparamspec),
 * qg_diagnostics.py:6-70 (entspec ... APEgenspec), qg_model.py:213-263
 * (APEflux, KEflux, APEgen; nz = 2) and layered_model.py:242-301
 * (KEspec_modal ... APEflux_div).  Spectral diagnostics have shape
 * [slots][nl][nk] per member (slots = nz for Ensspec, KEspec, KEspec_modal;
 * nz-1 for PEspec_modal; else 1); EKE has nz values, EKEdiss and APEgen one. */
enum {
  QGK_D_ENSSPEC = 0, QGK_D_KESPEC, QGK_D_EKEDISS, QGK_D_KEFRICTIONSPEC, QGK_D_EKE,
  QGK_D_DISSSPEC, QGK_D_ENSDISSSPEC, QGK_D_PARAMSPEC, QGK_D_ENSPARAMSPEC,
  QGK_D_ENTSPEC, QGK_D_PARAMSPEC_APEFLUX, QGK_D_PARAMSPEC_KEFLUX, QGK_D_ENSFLUX,
  QGK_D_ENSGENSPEC, QGK_D_ENSFRICTIONSPEC, QGK_D_APEGENSPEC,
  QGK_D_APEFLUX, QGK_D_KEFLUX, QGK_D_APEGEN,
  QGK_D_KESPEC_MODAL, QGK_D_PESPEC_MODAL, QGK_D_APESPEC, QGK_D_KEFLUX_DIV, QGK_D_APEFLUX_DIV,
  QGK_NDIAG
};
/* Per-member constants of the diagnostics, 2*nz*nz + 5*nz + 5 doubles:
 * Hi[nz], H, S[nz][nz] (stretching matrix), del1, del2, rd^-2, U1-U2,
 * Vbg[nz], Qx[nz], f2gpi[nz] (nz-1 used), inverse of the vertical-mode
 * matrix pmodes [nz][nz], kdi2[nz] (pyqg/model.py:284-313). */
int qgk_diag_set_params(qgk_bucket* b, int member, const double* params);
/* Adds every diagnostic whose bit is set in mask, from the current fields
 * (the point of _calc_diagnostics in model.py:406-436: after inversion,
 * advection, friction and parameterizations, before the time step).  dt and
 * the AB weights dt1..dt3 of the coming step enter Dissspec / ENSDissspec.
 * instant != 0 writes a one-shot evaluation to a separate set instead.
 * Needs QGK_FULL_STATE. */
int qgk_diag_accumulate(qgk_bucket* b, unsigned int mask, int instant, double dt, double dt1,
                        double dt2, double dt3);
/* Copies one diagnostic out: member >= 0 one member, -1 every member
 * ([M][slots][...]), -2 the sum over members.  Sums, not means: the caller
 * keeps the count. */
int qgk_diag_read(qgk_bucket* b, int diag, int member, int instant, double* host);
int qgk_diag_reset(qgk_bucket* b);

/* Built-in closures evaluated on the device from the current u, v, ph
 * (pyqg/parameterizations.py:132-289: Smagorinsky, ZannaBolton2020,
 * BackscatterBiharmonic).  The forcing times weight replaces (accumulate = 0)
 * or is added to the handle's du, dv (velocity closures) or dq (PV closure),
 * so that sums and multiples of closures (parameterizations.py:83-117) build
 * up in place; qgk_apply_uv_param / qgk_apply_q_param then add the fields to
 * dqhdt as kernel.pyx:397-442 does with host arrays.  With du_host/dv_host
 * (dq_host) given, the forcing is copied there instead and no field changes.
 * dx: grid spacing per member [M]; Hi: layer thicknesses [M][nz]. */
int qgk_param_smagorinsky(qgk_bucket* b, const double* dx, double constant, double weight,
                          int accumulate, double* du_host, double* dv_host);
int qgk_smagorinsky_viscosity(qgk_bucket* b, const double* dx, double constant,
                              double* nu_host);
int qgk_param_zb2020(qgk_bucket* b, double constant, double weight, int accumulate,
                     double* du_host, double* dv_host);
int qgk_param_backscatter(qgk_bucket* b, const double* dx, const double* Hi,
                          double smag_constant, double back_constant, double eps, double weight,
                          int accumulate, double* dq_host);
/* dqhdt += rfft2 of the handle's dq, or -il rfft2(du) + ik rfft2(dv)
 * (the device half of qgk_add_q_param / qgk_add_uv_param). */
int qgk_apply_q_param(qgk_bucket* b);
int qgk_apply_uv_param(qgk_bucket* b);

#ifdef __cplusplus
}
#endif
#endif
