// libqgk: MI355X (gfx950) FP64 time-stepping core for pyqg's PseudoSpectralKernel.
//
// Data flow of one step (pyqg/model.py:406-436 composed with kernel.pyx:312-526):
//   spectral qh --(inversion, velocities)--> ph, uh, vh
//   --(2-D c2r: y pass on columns, x pass on rows)--> u, v, q real
//   --(flux)--> uq = (u+Ubg) q, vq = v q
//   --(2-D r2c: x pass on rows, y pass on columns)--> uqh, vqh
//   --> dqhdt = -(ik uqh + il vqh + ikQy ph) [+ rek k2l2 ph bottom]
//   --> qh = filtr (qh + dt1 dqhdt + dt2 dqhdt_p + dt3 dqhdt_pp), history shift.
//
// Two execution paths share the same device code:
//  * per-phase kernels mirroring each reference method, keeping every
//    intermediate field the reference exposes (the drop-in path);
//  * a fused step of two kernels: a row kernel (x c2r of q,u,v -> flux -> x r2c)
//    and a column kernel (y FFT of the fluxes -> tendency -> drag -> AB3 +
//    filter -> inversion of the new qh -> y IFFT of q,u,v for the next step).
//    Between them only the x-mixed ("half transformed") fields travel through
//    HBM; ph, u, v, uq, vq and the real q never leave the chip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <utility>
#include <string>
#include <vector>

#include "fft_block.hpp"
#include "../../include/pyqg_amd.h"

using qgk::cd;
using qgk::cadd;
using qgk::csub;
using qgk::cmul;
using qgk::cscale;
using qgk::cconj;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPC(expr)                                                                  \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(QGK_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                              \
  } while (0)

// Per-member constants seen by every kernel.  Table pointers may be shared
// between members.  a == nullptr selects the closed-form two-layer inversion.
struct MemberParams {
  const double* kk;     // [nk]
  const double* ll;     // [nl]
  const double* k2l2;   // [nl][nk]
  const double* filtr;  // [nl][nk]
  const double* a;      // [nz][nz][nl][nk] or nullptr
  const double* Ubg;    // [nz]
  const double* Qy;     // [nz]
  double F1, F2;        // closed-form two-layer inversion
  double rek;
  double pad;
  // [3][nz] bands of a tridiagonal stretching matrix S (below, on and above
  // the diagonal) or nullptr: the fused nz >= 3 steps then solve
  // (S - kappa^2 I) ph = qh instead of reading the table a
  const double* tri;
  // ll[1]: with a standard grid (ll[l] = dl * l' for l' = l or l - nl, and
  // k2l2 = kk^2 + ll^2 exactly, checked on the device) the packed register
  // kernels form ll and kappa^2 from kk and dl instead of loading them
  double dl;
};

struct StepCoeffs {
  double dt1, dt2, dt3;
};

// Loads of the member constants through global-address-space pointers: a
// pointer read from MemberParams is generic, and generic (flat) loads count
// against both the memory and the LDS wait counters, so every LDS wait of the
// transforms would also wait for them.
__device__ __forceinline__ double gld(const double* p, size_t i) {
  return ((const __attribute__((address_space(1))) double*)p)[i];
}

// --------------------------------------------------------------------------
// launch geometry per transform length
// --------------------------------------------------------------------------
// Register radix of a transform of length N = 2^a or 3 * 2^a: lengths with the
// factor 3 need E divisible by 3 (6 or 12; a radix-24 pass would need 15-degree
// constants).
template <int N, int E2>
struct RadixFor {
  static constexpr int value =
      N % 3 == 0 ? (E2 <= 4 || N <= 48 ? 6 : 12) : (E2 < N ? E2 : N);
};

template <int N>
struct Geo {
  static constexpr int E = N % 3 == 0 ? RadixFor<N, 8>::value
                                      : (N <= 256 ? (N < 8 ? N : 8) : 16);  // elements per thread
  static constexpr int T = N / E;                            // threads per transform
  static constexpr int TK = N >= 1024 ? 4 : 8;               // columns per column tile
  static constexpr int RPraw = 128 / T;
  static constexpr int RP = RPraw < 1 ? 1 : RPraw;           // row pairs per row block
};

// Row stride of spectral arrays: nx/2+1 complex values padded to a multiple of
// 8 (128 bytes), so every row starts on a cache line and column tiles of 8
// read whole lines.
__host__ __device__ constexpr int ldk_of(int nx) { return ((nx / 2 + 1) + 7) / 8 * 8; }

// inversion at one spectral point: ph[k2] = sum_k1 a[k2,k1] qh[k1] (kernel.pyx:331-338)
__device__ __forceinline__ cd invert_point(const MemberParams& P, int nz, int k2,
                                           const cd* qh_pt, size_t plane, size_t pt) {
#pragma clang fp contract(off)
  // qh_pt: pointer to qh[0] at this point; layer stride = plane
  if (P.a) {
    cd acc = {0.0, 0.0};
    for (int k1 = 0; k1 < nz; ++k1) {
      double a = P.a[((size_t)k2 * nz + k1) * plane + pt];
      cd q = qh_pt[k1 * plane];
      // complex(a, 0) * q, as the reference multiplies by a complex matrix
      acc.x = acc.x + a * q.x;
      acc.y = acc.y + a * q.y;
    }
    return acc;
  }
  // two-layer closed form (qg_model.py:152-161); zero at kappa = 0
  double w = P.k2l2[pt];
  double det = w * (w + P.F1 + P.F2);
  if (det == 0.0) return {0.0, 0.0};
  double di = 1.0 / det;
  double a0, a1;
  if (k2 == 0) {
    a0 = -(w + P.F2) * di;
    a1 = -P.F1 * di;
  } else {
    a0 = -P.F2 * di;
    a1 = -(w + P.F1) * di;
  }
  cd q0 = qh_pt[0], q1 = qh_pt[plane];
  cd acc;
  acc.x = 0.0 + a0 * q0.x;
  acc.y = 0.0 + a0 * q0.y;
  acc.x = acc.x + a1 * q1.x;
  acc.y = acc.y + a1 * q1.y;
  return acc;
}

// Up to kPlaneJobs fields per transform launch: field f of launch block row
// blockIdx.y = f * nplanes + p is plane p of jobs.in[f] / jobs.out[f].
constexpr int kPlaneMax = 64;
constexpr int kPlaneJobs = 4;

struct PlaneJobs {
  const void* in[kPlaneJobs];
  void* out[kPlaneJobs];
};

// ==========================================================================
// Column (y) transforms of spectral planes [nl][nk]
// ==========================================================================
template <int NY, int SIGN>
__global__ __launch_bounds__(Geo<NY>::T * Geo<NY>::TK) void k_col_fft(
    PlaneJobs jobs, int nplanes, int nk, int ldk, const cd* __restrict__ tw) {
  using G = Geo<NY>;
  __shared__ cd lds[NY * G::TK];
  const int f = blockIdx.y / nplanes, pl = blockIdx.y - f * nplanes;
  const cd* in = static_cast<const cd*>(jobs.in[f]);
  cd* out = static_cast<cd*>(jobs.out[f]);
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int k = blockIdx.x * G::TK + c;
  const size_t plane = (size_t)pl * NY * ldk;
  const bool ok = k < nk;
  cd x[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s)
    x[s] = ok ? in[plane + (size_t)(t + s * G::T) * ldk + k] : cd{0.0, 0.0};
  qgk::fft_block<NY, G::E, SIGN, G::TK>(x, lds, t, c, tw);
  if (ok) {
#pragma unroll
    for (int s = 0; s < G::E; ++s) out[plane + (size_t)(t + s * G::T) * ldk + k] = x[s];
  }
}

// ==========================================================================
// Row (x) real transforms.  Two rows are carried by one complex transform:
// z = x_j + i x_{j+1}.
// ==========================================================================

// Hermitian extension of a half spectrum row h[0..N/2] at position p, with the
// imaginary parts of bins 0 and N/2 discarded (numpy / pocketfft c2r semantics).
// PK: the row is in the packed x-mixed form of the fused register path (slot 0
// holds bin 0 in its real part and bin N/2 in its imaginary part; N/2 slots).
template <int NX, bool PK = false>
__device__ __forceinline__ cd hext(const cd* __restrict__ h, int p) {
  cd v;
  if (PK && p == NX / 2) {
    v = {h[0].y, 0.0};
  } else if (p <= NX / 2) {
    v = h[p];
    if (p == 0 || p == NX / 2) v.y = 0.0;
  } else {
    v = h[NX - p];
    v.y = -v.y;
  }
  return v;
}

// c2r of planes [ny][nk] -> [ny][nx], scaled by `scale` (PK: packed x-mixed rows)
template <int NX, bool PK = false>
__global__ __launch_bounds__(Geo<NX>::T * Geo<NX>::RP) void k_row_c2r(
    PlaneJobs jobs, int nplanes, int ny, double scale, const cd* __restrict__ tw) {
  using G = Geo<NX>;
  constexpr int NK = ldk_of(NX);  // row stride
  __shared__ cd lds[NX * G::RP];
  const int f = blockIdx.y / nplanes, pl = blockIdx.y - f * nplanes;
  const cd* __restrict__ in = static_cast<const cd*>(jobs.in[f]);
  double* __restrict__ out = static_cast<double*>(jobs.out[f]);
  const int t = threadIdx.x % G::T;
  const int rp = threadIdx.x / G::T;
  const int jp = blockIdx.x * G::RP + rp;
  const bool ok = 2 * jp < ny;
  const int j0 = ok ? 2 * jp : 0;
  const cd* r0 = in + ((size_t)pl * ny + j0) * NK;
  const cd* r1 = r0 + NK;
  cd x[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const int p = t + s * G::T;
    cd a = hext<NX, PK>(r0, p), b = hext<NX, PK>(r1, p);
    x[s] = {a.x - b.y, a.y + b.x};
  }
  qgk::fft_block_rows<NX, G::E, +1>(x, lds, t, rp, tw);
  if (ok) {
    double* o0 = out + ((size_t)pl * ny + j0) * NX;
    double* o1 = o0 + NX;
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const int p = t + s * G::T;
      o0[p] = x[s].x * scale;
      o1[p] = x[s].y * scale;
    }
  }
}

// Splits the transform Z of z = x0 + i x1 (held in LDS, one swizzled region
// per row pair) into
// the half spectra of x0 and x1 and writes them.
template <int NX, int RP>
__device__ __forceinline__ void split_store(const cd* lds, int t, int rp, bool ok, cd* o0,
                                            cd* o1) {
  using G = Geo<NX>;
  if (!ok) return;
  for (int k = t; k <= NX / 2; k += G::T) {
    cd A = lds[qgk::RowSwizzled<NX>::at(k, rp)];
    cd B = lds[qgk::RowSwizzled<NX>::at(k == 0 ? 0 : NX - k, rp)];
    o0[k] = {0.5 * (A.x + B.x), 0.5 * (A.y - B.y)};
    o1[k] = {0.5 * (A.y + B.y), -0.5 * (A.x - B.x)};
  }
}

// r2c of planes [ny][nx] -> [ny][nk]
template <int NX>
__global__ __launch_bounds__(Geo<NX>::T * Geo<NX>::RP) void k_row_r2c(
    PlaneJobs jobs, int nplanes, int ny, const cd* __restrict__ tw) {
  using G = Geo<NX>;
  constexpr int NK = ldk_of(NX);  // row stride
  __shared__ cd lds[NX * G::RP];
  const int f = blockIdx.y / nplanes, pl = blockIdx.y - f * nplanes;
  const double* __restrict__ in = static_cast<const double*>(jobs.in[f]);
  cd* __restrict__ out = static_cast<cd*>(jobs.out[f]);
  const int t = threadIdx.x % G::T;
  const int rp = threadIdx.x / G::T;
  const int jp = blockIdx.x * G::RP + rp;
  const bool ok = 2 * jp < ny;
  const int j0 = ok ? 2 * jp : 0;
  const double* i0 = in + ((size_t)pl * ny + j0) * NX;
  const double* i1 = i0 + NX;
  cd x[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const int p = t + s * G::T;
    x[s] = {i0[p], i1[p]};
  }
  qgk::fft_block_rows<NX, G::E, -1>(x, lds, t, rp, tw);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G::E; ++s) lds[qgk::RowSwizzled<NX>::at(t, s * G::T, rp)] = x[s];
  __syncthreads();
  cd* o0 = out + ((size_t)pl * ny + j0) * NK;
  split_store<NX, G::RP>(lds, t, rp, ok, o0, o0 + NK);
}

// ==========================================================================
// One-launch 2-D transforms of small square planes (n <= 64).  A small model
// steps through dozens of short launches per step, each bound by launch
// latency rather than by its work, so here one workgroup transforms a whole
// plane: the row and column passes of k_row_r2c / k_row_c2r and k_col_fft (the
// same Stockham passes, radix and twiddles) with the plane exchanged through
// LDS instead of through global memory.  One LDS region of N x (N/2 + 1)
// values serves in turn as the row transforms' exchange, the half-spectrum
// tile and the column transforms' exchange: every transform's first LDS write
// follows a barrier, after the reads of the previous use.  A launch carries up
// to kPlaneJobs fields of nplanes planes each.
// ==========================================================================

// T threads per transform; N/2 + 1 groups of them: the columns of the half
// spectrum in the column pass, the N/2 row pairs (and one idle group) in the
// row pass.
template <int N>
struct PlaneGeo {
  static constexpr int E = RadixFor<N, 4>::value;
  static constexpr int T = N / E;
  static constexpr int NK = N / 2 + 1;
  static constexpr int THREADS = T * NK;
};

// rfftn of real planes [N][N] -> half spectra [N][ldk] (padding columns zeroed)
template <int N>
__global__ __launch_bounds__(PlaneGeo<N>::THREADS) void k_plane_r2c(
    PlaneJobs jobs, int nplanes, int ldk, const cd* __restrict__ twx,
    const cd* __restrict__ twy) {
  using G = PlaneGeo<N>;
  using RS = qgk::RowSwizzled<N>;
  __shared__ cd lds[N * G::NK];
  const int f = blockIdx.x / nplanes, p = blockIdx.x - f * nplanes;
  const double* in = static_cast<const double*>(jobs.in[f]) + (size_t)p * N * N;
  cd* out = static_cast<cd*>(jobs.out[f]) + (size_t)p * N * ldk;
  const int t = threadIdx.x % G::T;
  const int g = threadIdx.x / G::T;
  // row pass: group g carries rows 2g and 2g + 1 as z = x_2g + i x_2g+1
  const int j0 = g < N / 2 ? 2 * g : 0;
  cd x[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const int q = t + s * G::T;
    x[s] = {in[j0 * N + q], in[(j0 + 1) * N + q]};
  }
  qgk::fft_block_rows<N, G::E, -1>(x, lds, t, g, twx);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G::E; ++s) lds[RS::at(t, s * G::T, g)] = x[s];
  __syncthreads();
  // column pass: group g is column k = g; row r is split out of pair r / 2
  const int k = g, km = k == 0 ? 0 : N - k;
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const int r = t + s * G::T;
    const cd A = lds[RS::at(k, r >> 1)], B = lds[RS::at(km, r >> 1)];
    x[s] = (r & 1) ? cd{0.5 * (A.y + B.y), -0.5 * (A.x - B.x)}
                   : cd{0.5 * (A.x + B.x), 0.5 * (A.y - B.y)};
  }
  qgk::fft_block<N, G::E, -1, G::NK>(x, lds, t, k, twy);
#pragma unroll
  for (int s = 0; s < G::E; ++s) out[(size_t)(t + s * G::T) * ldk + k] = x[s];
  if (G::NK + g < ldk) {
#pragma unroll
    for (int s = 0; s < G::E; ++s)
      out[(size_t)(t + s * G::T) * ldk + G::NK + g] = cd{0.0, 0.0};
  }
}

// Spectral operators applied on the way into an inverse transform, per field:
//   kSpecPlain: X itself;
//   kSpecIkil:  cx (i k) X + cy (i l) Y, X or Y possibly absent, the products
//               and sums in k_spec_ikil's order;
//   kSpecWv2:   cx (kappa^2)^power X, as k_spec_wv2 forms it (cx the sign).
enum { kSpecPlain = 0, kSpecIkil = 1, kSpecWv2 = 2 };
struct SpecJobs {
  const cd* X[kPlaneJobs];
  const cd* Y[kPlaneJobs];
  double cx[kPlaneJobs], cy[kPlaneJobs];
  int kind[kPlaneJobs], power[kPlaneJobs];
};

// irfftn of half spectra [N][ldk] -> real planes [N][N], scaled by `scale`
// (numpy c2r: the imaginary parts of bins kx = 0 and N/2 are dropped).  With
// SPEC the half spectra are the operator results of `dj` (planes of member
// p / nz) instead of jobs.in.
template <int N, bool SPEC>
__global__ __launch_bounds__(PlaneGeo<N>::THREADS) void k_plane_c2r(
    PlaneJobs jobs, SpecJobs dj, const MemberParams* __restrict__ mp, int nz, int nplanes,
    int ldk, double scale, const cd* __restrict__ twx, const cd* __restrict__ twy) {
  using G = PlaneGeo<N>;
  __shared__ cd lds[N * G::NK];
  const int f = blockIdx.x / nplanes, p = blockIdx.x - f * nplanes;
  double* out = static_cast<double*>(jobs.out[f]) + (size_t)p * N * N;
  const int t = threadIdx.x % G::T;
  const int g = threadIdx.x / G::T;
  // column pass: group g is column k = g
  cd x[G::E];
  if constexpr (SPEC) {
    const size_t base = (size_t)p * N * ldk;
    const MemberParams& P = mp[p / nz];
    const cd* X = dj.X[f];
    const cd* Y = dj.Y[f];
    const double cx = dj.cx[f], cy = dj.cy[f];
    const int kind = dj.kind[f], power = dj.power[f];
    const double kk = P.kk[g];
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const int l = t + s * G::T;
      const size_t pt = (size_t)l * ldk + g;
      const size_t i = base + pt;
      cd r = {0.0, 0.0};
      if (kind == kSpecPlain) {
        r = X[i];
      } else if (kind == kSpecIkil) {
        if (X) {
          const cd v = X[i];
          r = {-kk * v.y, kk * v.x};
          r = {cx * r.x, cx * r.y};
        }
        if (Y) {
          const double ll = P.ll[l];
          const cd v = Y[i];
          r = {r.x + cy * (-ll * v.y), r.y + cy * (ll * v.x)};
        }
      } else {
        const double w = P.k2l2[pt];
        double fw = cx;
        for (int q = 0; q < power; ++q) fw *= w;
        const cd v = X[i];
        r = {fw * v.x, fw * v.y};
      }
      x[s] = r;
    }
  } else {
    const cd* in = static_cast<const cd*>(jobs.in[f]) + (size_t)p * N * ldk;
#pragma unroll
    for (int s = 0; s < G::E; ++s) x[s] = in[(size_t)(t + s * G::T) * ldk + g];
  }
  qgk::fft_block<N, G::E, +1, G::NK>(x, lds, t, g, twy);
  __syncthreads();
  // the y-transformed half spectra, row r at lds + r * NK
#pragma unroll
  for (int s = 0; s < G::E; ++s) lds[(t + s * G::T) * G::NK + g] = x[s];
  __syncthreads();
  // row pass: group g carries rows 2g and 2g + 1 (group N/2 idles)
  const bool ok = g < N / 2;
  const int j0 = ok ? 2 * g : 0;
  const cd* r0 = lds + j0 * G::NK;
  const cd* r1 = r0 + G::NK;
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const int q = t + s * G::T;
    const cd a = hext<N>(r0, q), c = hext<N>(r1, q);
    x[s] = {a.x - c.y, a.y + c.x};
  }
  qgk::fft_block_rows<N, G::E, +1>(x, lds, t, g, twx);
  if (ok) {
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const int q = t + s * G::T;
      out[j0 * N + q] = x[s].x * scale;
      out[(j0 + 1) * N + q] = x[s].y * scale;
    }
  }
}

// ==========================================================================
// Pointwise phases (drop-in path).  Contraction is disabled so every product
// and sum rounds exactly as the reference's loops do.
// ==========================================================================

// _invert, spectral part (kernel.pyx:323-347): ph, uh = -il ph, vh = ik ph
__global__ void k_invert(const cd* __restrict__ qh, cd* __restrict__ ph, cd* __restrict__ uh,
                         cd* __restrict__ vh, const MemberParams* __restrict__ params, int nz,
                         int nl, int nk, int M) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)nl * nk;
  const size_t total = (size_t)M * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(idx / plane);
    const size_t pt = idx - (size_t)m * plane;
    const int l = (int)(pt / nk), k = (int)(pt - (size_t)l * nk);
    const MemberParams& P = params[m];
    const size_t base = (size_t)m * nz * plane + pt;
    const double kk = P.kk[k], ll = P.ll[l];
    for (int z = 0; z < nz; ++z) {
      cd p = invert_point(P, nz, z, qh + (size_t)m * nz * plane + pt, plane, pt);
      ph[base + z * plane] = p;
      // -(1j*ll) * p and (1j*kk) * p
      uh[base + z * plane] = {ll * p.y, -ll * p.x};
      vh[base + z * plane] = {-kk * p.y, kk * p.x};
    }
  }
}

// _do_advection flux (kernel.pyx:372-378): uq = (u + Ubg) q, vq = v q
__global__ void k_flux(const double* __restrict__ u, const double* __restrict__ v,
                       const double* __restrict__ q, double* __restrict__ uq,
                       double* __restrict__ vq, const MemberParams* __restrict__ params, int nz,
                       size_t plane, int M) {
#pragma clang fp contract(off)
  const size_t total = (size_t)M * nz * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t pl = idx / plane;
    const int m = (int)(pl / nz), z = (int)(pl - (size_t)m * nz);
    const double qq = q[idx];
    uq[idx] = (u[idx] + params[m].Ubg[z]) * qq;
    vq[idx] = v[idx] * qq;
  }
}

// _do_advection divergence (kernel.pyx:386-394):
// dqhdt = -(ik uqh + il vqh + ikQy ph)
__global__ void k_divergence(const cd* __restrict__ uqh, const cd* __restrict__ vqh,
                             const cd* __restrict__ ph, cd* __restrict__ dqhdt,
                             const MemberParams* __restrict__ params, int nz, int nl, int nk,
                             int M) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)nl * nk;
  const size_t total = (size_t)M * nz * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t pl = idx / plane;
    const size_t pt = idx - pl * plane;
    const int m = (int)(pl / nz), z = (int)(pl - (size_t)m * nz);
    const int l = (int)(pt / nk), k = (int)(pt - (size_t)l * nk);
    const MemberParams& P = params[m];
    const double kk = P.kk[k], ll = P.ll[l];
    const double kq = kk * P.Qy[z];
    const cd a = uqh[idx], b = vqh[idx], p = ph[idx];
    cd s = {-kk * a.y + -ll * b.y, kk * a.x + ll * b.x};
    s = {s.x + -kq * p.y, s.y + kq * p.x};
    dqhdt[idx] = {-s.x, -s.y};
  }
}

// _do_friction (kernel.pyx:444-461): bottom layer, dqhdt += (rek k2l2) ph
__global__ void k_friction(const cd* __restrict__ ph, cd* __restrict__ dqhdt,
                           const MemberParams* __restrict__ params, int nz, int nl, int nk,
                           int M) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)nl * nk;
  const size_t total = (size_t)M * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(idx / plane);
    const size_t pt = idx - (size_t)m * plane;
    const MemberParams& P = params[m];
    if (P.rek == 0.0) continue;
    const size_t o = ((size_t)m * nz + nz - 1) * plane + pt;
    const double c = P.rek * P.k2l2[pt];
    const cd p = ph[o];
    cd d = dqhdt[o];
    dqhdt[o] = {d.x + c * p.x, d.y + c * p.y};
  }
}

// q parameterization add (kernel.pyx:435-441): dqhdt += dqh
__global__ void k_add(cd* __restrict__ dqhdt, const cd* __restrict__ dqh, size_t total) {
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    cd d = dqhdt[idx], e = dqh[idx];
    dqhdt[idx] = {d.x + e.x, d.y + e.y};
  }
}

// Stochastic forcing: a first-order autoregressive process per spectral point,
//   xi <- phi xi + sqrt(1 - phi^2) amp w,   dqh += xi,
// w complex normal (E|w|^2 = 1) from a counter-based hash of (seed, step
// counter, point), so that a captured step draws new numbers at every replay
// (the counter lives in device memory; k_noise_count advances it after each
// draw).  The columns kx = 0 and kx = nx/2 of a real field's half spectrum
// are Hermitian in l: the draw at row ny - l is the conjugate of row l's (the
// self-conjugate rows draw a real w of unit variance), so xi stays the
// spectrum of a real field as long as amp is symmetric in l there.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__global__ void k_spectral_noise(cd* __restrict__ dqh, cd* __restrict__ xi,
                                 const double* __restrict__ amp, double phi, double sig,
                                 unsigned long long seed, const unsigned long long* __restrict__ ctr,
                                 int M, int nz, int ny, int nx, int nk, int ldk) {
  const unsigned long long step = *ctr;
  const size_t total = (size_t)M * nz * ny * nk;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % nk);
    const int l = (int)((i / nk) % ny);
    const size_t zm = i / ((size_t)nk * ny);  // member * nz + layer
    const int z = (int)(zm % nz);
    const bool herm = k == 0 || (nx % 2 == 0 && k == nx / 2);
    const int lm = l == 0 ? 0 : ny - l;
    const int lc = herm && lm < l ? lm : l;  // the row whose draw this point uses
    const unsigned long long key =
        mix64(seed ^ mix64(step * 0x100000001B3ull + ((zm * ny + lc) * nk + k)));
    const double u1 = ((key >> 11) + 1) * (1.0 / 9007199254740993.0);   // (0, 1]
    const double u2 = (mix64(key) >> 11) * (1.0 / 9007199254740992.0);  // [0, 1)
    const double r = sqrt(-log(u1));  // |w|^2 exponential with mean 1
    double s, c;
    sincos(6.283185307179586 * u2, &s, &c);
    cd w = {r * c, r * s};
    if (herm) {
      if (lc == lm && lm == l) w = {w.x * 1.4142135623730951, 0.0};  // self-conjugate row
      else if (lc != l) w.y = -w.y;                                  // conjugate of row lc
    }
    const double a = amp[((size_t)z * ny + l) * nk + k];
    const size_t o = (zm * ny + l) * ldk + k;
    const cd x = xi[o];
    const cd xn = {phi * x.x + sig * a * w.x, phi * x.y + sig * a * w.y};
    xi[o] = xn;
    const cd d = dqh[o];
    dqh[o] = {d.x + xn.x, d.y + xn.y};
  }
}

__global__ void k_noise_count(unsigned long long* __restrict__ ctr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *ctr = *ctr + 1;
}

// uv parameterization add (kernel.pyx:412-421): dqhdt += (-il duh + ik dvh)
__global__ void k_add_curl(cd* __restrict__ dqhdt, const cd* __restrict__ duh,
                           const cd* __restrict__ dvh, const MemberParams* __restrict__ params,
                           int nz, int nl, int nk, int M) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)nl * nk;
  const size_t total = (size_t)M * nz * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t pl = idx / plane;
    const size_t pt = idx - pl * plane;
    const int m = (int)(pl / nz);
    const int l = (int)(pt / nk), k = (int)(pt - (size_t)l * nk);
    const double kk = params[m].kk[k], ll = params[m].ll[l];
    const cd a = duh[idx], b = dvh[idx];
    const cd c = {ll * a.y + -kk * b.y, -ll * a.x + kk * b.x};
    const cd d = dqhdt[idx];
    dqhdt[idx] = {d.x + c.x, d.y + c.y};
  }
}

// _forward_timestep spectral part (kernel.pyx:502-512).  The history rotates by
// pointer: the new dqhdt_p is written into the old dqhdt_pp buffer.
__global__ void k_ab_step(cd* __restrict__ qh, const cd* __restrict__ d,
                          const cd* __restrict__ p, cd* __restrict__ pp_to_newp,
                          const MemberParams* __restrict__ params, StepCoeffs c, int nz,
                          int nl, int nk, int M) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)nl * nk;
  const size_t total = (size_t)M * nz * plane;
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < total;
       idx += (size_t)gridDim.x * blockDim.x) {
    const size_t pl = idx / plane;
    const size_t pt = idx - pl * plane;
    const int m = (int)(pl / nz);
    const double f = params[m].filtr[pt];
    const cd q = qh[idx], dd = d[idx], pv = p[idx], ppv = pp_to_newp[idx];
    double rx = q.x + c.dt1 * dd.x;
    double ry = q.y + c.dt1 * dd.y;
    rx = rx + c.dt2 * pv.x;
    ry = ry + c.dt2 * pv.y;
    rx = rx + c.dt3 * ppv.x;
    ry = ry + c.dt3 * ppv.y;
    qh[idx] = {f * rx, f * ry};
    pp_to_newp[idx] = dd;
  }
}

// ==========================================================================
// Fused step
// ==========================================================================

// Launch geometry of the fused kernels for transform length N with E elements
// per thread: T = N/E threads per transform, TK columns per column tile, RP row
// pairs per row block.  E sets the register radix of the Stockham passes and is
// the autotuned parameter.
template <int N, int E_>
struct FGeo {
  static constexpr int E = E_ < N ? E_ : N;
  static constexpr int T = N / E;
  static constexpr int TKraw = 512 / T;  // 8 columns = one 128-byte line per row
  static constexpr int TK = TKraw < 1 ? 1 : (TKraw > 8 ? 8 : TKraw);
  static constexpr int RPraw = 128 / T;
  static constexpr int RP = RPraw < 1 ? 1 : RPraw;
};

// Transform instantiation tags of the fused kernels (see fft_block): each hot
// kernel compiles its own copy of the transforms.
constexpr int kTagRow = 1, kTagColReg = 2, kTagColSplit = 3, kTagColPh = 4;

// Streaming accesses.  The fused step touches every spectral field once per
// step, so its loads (and the column kernels' stores) carry the non-temporal
// hint: the L2 is left to the small tables every block re-reads (filter,
// kappa^2, wavenumbers, twiddles) instead of being flushed through by the
// fields.  The row kernel's stores stay ordinary: with the hint they measured
// 5 % slower and read as 14 % more written bytes.
__device__ __forceinline__ cd ld_stream(const cd* p) {
  return {__builtin_nontemporal_load(&p->x), __builtin_nontemporal_load(&p->y)};
}
__device__ __forceinline__ void st_stream(cd* p, cd v) {
  __builtin_nontemporal_store(v.x, &p->x);
  __builtin_nontemporal_store(v.y, &p->y);
}

// The twiddle table of a length-N transform copied into LDS at block start.
// Every Stockham pass after the first reads one twiddle right after a
// barrier; from LDS that read costs an LDS round trip instead of an L2 one on
// the pass's critical path.  No barrier is needed here: the first pass reads
// no twiddle and each later pass begins with a barrier.
template <int N>
__device__ __forceinline__ const cd* stage_twiddles(cd* twl, const cd* __restrict__ tw) {
  for (int i = threadIdx.x; i < N; i += blockDim.x) twl[i] = tw[i];
  return twl;
}

// Splits the transform Z of a row pair (in LDS) into the rows' half spectra
// and stores them.  Unpacked: slots 0..NX/2 and zeros in the padding up to the
// row stride, so whole 128-byte lines are written.  Packed (PK): slots
// 0..NX/2-1, slot 0 holding the real bins 0 and NX/2 as one complex value.
template <int NX, int T, int RP, bool PK = false>
__device__ __forceinline__ void split_store_t(const cd* lds, int t, int rp, bool ok, cd* o0,
                                              cd* o1) {
  if (!ok) return;
  if constexpr (PK) {
    for (int k = t; k < NX / 2; k += T) {
      cd A = lds[qgk::RowSwizzled<NX>::at(k, rp)];
      if (k == 0) {
        // Z(0) and Z(NX/2) are real for each row: re -> row j, im -> row j+1
        const cd H = lds[qgk::RowSwizzled<NX>::at(NX / 2, rp)];
        o0[0] = {A.x, H.x};
        o1[0] = {A.y, H.y};
      } else {
        cd B = lds[qgk::RowSwizzled<NX>::at(NX - k, rp)];
        o0[k] = {0.5 * (A.x + B.x), 0.5 * (A.y - B.y)};
        o1[k] = {0.5 * (A.y + B.y), -0.5 * (A.x - B.x)};
      }
    }
    return;
  }
  constexpr int NK = ldk_of(NX);
  for (int k = t; k < NK; k += T) {
    if (k <= NX / 2) {
      cd A = lds[qgk::RowSwizzled<NX>::at(k, rp)];
      cd B = lds[qgk::RowSwizzled<NX>::at(k == 0 ? 0 : NX - k, rp)];
      o0[k] = {0.5 * (A.x + B.x), 0.5 * (A.y - B.y)};
      o1[k] = {0.5 * (A.y + B.y), -0.5 * (A.x - B.x)};
    } else {  // padding: zeros, so the row is written as whole 128-byte lines
      o0[k] = cd{0.0, 0.0};
      o1[k] = cd{0.0, 0.0};
    }
  }
}

// Hermitian extension of a row pair for one complex c2r transform:
// z[s] = X0(p) + i X1(p), p = t + s*T, where X_r(p) = h_r[p] for p <= NX/2 and
// conj(h_r[NX - p]) above, with the imaginary parts of bins 0 and NX/2
// dropped (numpy's c2r).  h holds the thread's own slots p < NX/2 of both rows;
// they are exchanged through LDS (rows at [0, NX/2) and [NX/2, NX) of each
// lane) for the mirrored positions; the Nyquist bin is read by thread 0 (PK:
// it is the imaginary part of the packed slot 0 thread 0 already holds).
template <int NX, int E, int T, int RP, bool PK = false>
__device__ __forceinline__ void row_pair_extend(const cd (&h)[2][E / 2],
                                                const cd* __restrict__ row, int NK,
                                                cd (&z)[E], cd* lds, int t, int rp) {
  constexpr int HALF = NX / 2;
  __syncthreads();  // the previous transform's last LDS reads are done
#pragma unroll
  for (int s = 0; s < E / 2; ++s) {
    lds[qgk::RowSwizzled<NX>::at(t, s * T, rp)] = h[0][s];
    lds[qgk::RowSwizzled<NX>::at(t, HALF + s * T, rp)] = h[1][s];
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < E; ++s) {
    const int p = t + s * T;
    cd a, b;
    if (s < E / 2) {
      a = h[0][s];
      b = h[1][s];
      if (p == 0) {
        a.y = 0.0;
        b.y = 0.0;
      }
    } else if (p == HALF) {
      if constexpr (PK) {
        a = {h[0][0].y, 0.0};
        b = {h[1][0].y, 0.0};
      } else {
        a = ld_stream(row + HALF);
        b = ld_stream(row + NK + HALF);
        a.y = 0.0;
        b.y = 0.0;
      }
    } else {
      // element NX - p = -t + (NX - s*T)
      a = lds[qgk::RowSwizzled<NX>::at(-t, NX - s * T, rp)];
      b = lds[qgk::RowSwizzled<NX>::at(-t, HALF + NX - s * T, rp)];
      a.y = -a.y;
      b.y = -b.y;
    }
    z[s] = {a.x - b.y, a.y + b.x};
  }
}

// Row phase: per (member, layer, row pair): x-c2r of q, u, v, flux, x-r2c.
// In:  Qm, Um, Vm  [M][nz][ny][ldk] (y already inverse transformed)
// Out: UQm, VQm    [M][nz][ny][ldk] (x already forward transformed)
// PK selects the packed x-mixed rows of the register column kernels: NX/2
// slots per row, bins 0 and NX/2 (real after an x transform of real data, and
// only their real parts enter a c2r) sharing slot 0.  The row stride stays
// ldk, so a row is NX/2 * 16 bytes of whole lines instead of ldk * 16.
// Each complex transform carries the two rows of the pair (row j in the real
// part, row j+1 in the imaginary part), so the six input rows are loaded at
// once and the five transforms run back to back.
// Waves per SIMD the row kernel is compiled for: radix 8 at 3 (a few spilled
// registers) rather than the 2 its natural 172 VGPRs allow.
template <int E>
constexpr int row_flux_waves() { return E == 8 ? 3 : 1; }

// The BF16 rounding of k_pack_nhwc (nearest even), for the network input.
__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned int bits = __float_as_uint(f);
  bits += 0x7fffu + ((bits >> 16) & 1u);
  return (unsigned short)(bits >> 16);
}

// NET: the kernel also writes the learned closure's input from the real q, u,
// v it has formed, as qgk_pack_inputs does from the real fields: NHWC BF16,
// channel f * nz + z (f = q, u, v) of pixel (row, x) of the member is the
// value times net_scale[f * nz + z]; a block writes the three channels of its
// layer.  The padding channels are left as they are (zero in the caller's
// buffer).
template <int NX, int E, bool PK, bool NET = false>
__global__ __launch_bounds__((FGeo<NX, E>::T * FGeo<NX, E>::RP), row_flux_waves<E>()) void k_row_flux(
    const cd* __restrict__ Qm, const cd* __restrict__ Um, const cd* __restrict__ Vm,
    cd* __restrict__ UQm, cd* __restrict__ VQm, const MemberParams* __restrict__ params,
    int nz, int ny, double scale, const cd* __restrict__ tw,
    const double* __restrict__ net_scale = nullptr, unsigned short* __restrict__ net_in = nullptr,
    int cinp = 0) {
  using G = FGeo<NX, E>;
  constexpr int NK = ldk_of(NX);  // row stride
  __shared__ cd lds[NX * G::RP];
  __shared__ cd twl[NX];
  tw = stage_twiddles<NX>(twl, tw);
  const int t = threadIdx.x % G::T;
  const int rp = threadIdx.x / G::T;
  const int jp = blockIdx.x * G::RP + rp;
  const bool ok = 2 * jp < ny;
  const int j0 = ok ? 2 * jp : 0;
  const int plane_id = blockIdx.y;  // m * nz + z
  const int m = plane_id / nz, z = plane_id - m * nz;
  const double ubg = params[m].Ubg[z];
  const size_t row0 = ((size_t)plane_id * ny + j0) * NK;

  // Every half-spectrum element of the six rows is read from memory once: the
  // positions p < NX/2 a thread owns come straight from global memory, the
  // mirrored ones (p > NX/2, conj of NX - p) from the LDS copy of the pair.
  constexpr int H = G::E / 2;  // slots with p = t + s*T < NX/2
  cd hq[2][H], hu[2][H], hv[2][H];
#pragma unroll
  for (int s = 0; s < H; ++s) {
    const int p = t + s * G::T;
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      hq[r][s] = ld_stream(Qm + row0 + r * NK + p);
      hu[r][s] = ld_stream(Um + row0 + r * NK + p);
      hv[r][s] = ld_stream(Vm + row0 + r * NK + p);
    }
  }
  cd zq[G::E], zu[G::E], zv[G::E];
  row_pair_extend<NX, G::E, G::T, G::RP, PK>(hq, Qm + row0, NK, zq, lds, t, rp);
  qgk::fft_block_rows<NX, G::E, +1, kTagRow>(zq, lds, t, rp, tw);
  row_pair_extend<NX, G::E, G::T, G::RP, PK>(hu, Um + row0, NK, zu, lds, t, rp);
  qgk::fft_block_rows<NX, G::E, +1, kTagRow>(zu, lds, t, rp, tw);
  row_pair_extend<NX, G::E, G::T, G::RP, PK>(hv, Vm + row0, NK, zv, lds, t, rp);
  qgk::fft_block_rows<NX, G::E, +1, kTagRow>(zv, lds, t, rp, tw);
  // fluxes of both rows: uq = (u + Ubg) q, vq = v q (kernel.pyx:372-378)
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const double q0 = zq[s].x * scale, q1 = zq[s].y * scale;
    const double u0 = zu[s].x * scale, u1 = zu[s].y * scale;
    const double v0 = zv[s].x * scale, v1 = zv[s].y * scale;
    if constexpr (NET) {
      if (ok) {
        const size_t px0 = (((size_t)m * ny + j0) * NX + t + s * G::T) * cinp;
        const size_t px1 = px0 + (size_t)NX * cinp;  // row j0 + 1
        const double sq = net_scale[z], su = net_scale[nz + z], sv = net_scale[2 * nz + z];
        net_in[px0 + z] = bf16_rne((float)(q0 * sq));
        net_in[px1 + z] = bf16_rne((float)(q1 * sq));
        net_in[px0 + nz + z] = bf16_rne((float)(u0 * su));
        net_in[px1 + nz + z] = bf16_rne((float)(u1 * su));
        net_in[px0 + 2 * nz + z] = bf16_rne((float)(v0 * sv));
        net_in[px1 + 2 * nz + z] = bf16_rne((float)(v1 * sv));
      }
    }
    zu[s] = {(u0 + ubg) * q0, (u1 + ubg) * q1};
    zv[s] = {v0 * q0, v1 * q1};
  }
  qgk::fft_block_rows<NX, G::E, -1, kTagRow>(zu, lds, t, rp, tw);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G::E; ++s) lds[qgk::RowSwizzled<NX>::at(t, s * G::T, rp)] = zu[s];
  __syncthreads();
  split_store_t<NX, G::T, G::RP, PK>(lds, t, rp, ok, UQm + row0, UQm + row0 + NK);
  qgk::fft_block_rows<NX, G::E, -1, kTagRow>(zv, lds, t, rp, tw);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < G::E; ++s) lds[qgk::RowSwizzled<NX>::at(t, s * G::T, rp)] = zv[s];
  __syncthreads();
  split_store_t<NX, G::T, G::RP, PK>(lds, t, rp, ok, VQm + row0, VQm + row0 + NK);
}

// Standard-grid check of every member (block per member): ll[l] = dl * l'
// (l' = l below ny/2, l - ny above) and k2l2[l][k] = kk[k]^2 + ll[l]^2, both
// exactly, as the model's wavenumber tables are built (model.py:461-463,
// kernel.pyx:543-546).  The packed register kernels then form ll and kappa^2
// themselves instead of loading them; tables set to anything else keep the
// table-reading kernels.
__global__ __launch_bounds__(256) void k_grid_check(const MemberParams* __restrict__ params,
                                                    int ny, int nk, int ldk,
                                                    int* __restrict__ ok) {
#pragma clang fp contract(off)
  const MemberParams& P = params[blockIdx.x];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  int b = 0;
  for (int l = threadIdx.x; l < ny; l += blockDim.x) {
    const double ll = P.dl * (double)(l < ny / 2 ? l : l - ny);
    if (!(P.ll[l] == ll)) b = 1;
    for (int k = 0; k < nk; ++k) {
      const double kk = P.kk[k];
      if (!(P.k2l2[(size_t)l * ldk + k] == kk * kk + ll * ll)) b = 1;
    }
  }
  if (b) atomicOr(&bad, 1);
  __syncthreads();
  if (threadIdx.x == 0) ok[blockIdx.x] = bad ? 0 : 1;
}

// Health check of every member: the reference asserts CFL < 1 every twrite
// steps (model.py:564-574 with qg_diagnostics.py:76-79), CFL being
// max|u + Ubg|, |v| over the grid times dt/dx.  Here each member's largest
// speed comes from the primed x-mixed q, u, v, transformed back to real space
// one field at a time.  The maximum is kept as the bit pattern of a
// non-negative double, which orders like its value, so an unsigned atomic max
// builds it; |NaN| has a larger pattern than +inf, so a NaN in u or v, or a
// non-finite q anywhere in the member, leaves a NaN (or inf) in its word.
__device__ __forceinline__ unsigned long long abs_bits(double x) {
  return (unsigned long long)__double_as_longlong(x) & 0x7fffffffffffffffULL;
}

// One field of every member in real space (the x inverse transform of its
// x-mixed copy, [M][nz][ny][nx]) folded into the members' health words: mode 0
// adds the layer's Ubg (u), mode 1 takes the value (v), mode 2 only flags
// non-finite values (q).  Block per (part of a plane); a wave maximum, then
// one atomic per wave.
__global__ __launch_bounds__(256) void k_speed_fold(const double* __restrict__ r, int nplane,
                                                    int nz, const MemberParams* __restrict__ params,
                                                    int mode, unsigned long long* __restrict__ speed) {
  constexpr unsigned long long kInf = 0x7ff0000000000000ULL;
  const int plane_id = blockIdx.y;
  const int m = plane_id / nz, z = plane_id - m * nz;
  const double add = mode == 0 ? params[m].Ubg[z] : 0.0;
  const double* p = r + (size_t)plane_id * nplane;
  unsigned long long w = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nplane; i += gridDim.x * blockDim.x) {
    const unsigned long long b = abs_bits(p[i] + add);
    if (mode < 2 || b >= kInf) w = b > w ? b : w;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned lo = __shfl_xor((unsigned)(w & 0xffffffffULL), off);
    const unsigned hi = __shfl_xor((unsigned)(w >> 32), off);
    const unsigned long long o = ((unsigned long long)hi << 32) | lo;
    w = o > w ? o : w;
  }
  if ((threadIdx.x & 63) == 0 && w) atomicMax(speed + m, w);
}

// Scratch of the allocation-free health check (qgk_member_speed_noalloc): one
// health word per member in a file-scope device array, so the call needs no
// device allocation.  Host code uses its per-device address from
// hipGetSymbolAddress, never the array's name.
constexpr int kNoAllocMaxMembers = 4096;
static __device__ __align__(256) unsigned long long scratch[kNoAllocMaxMembers];

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_u64_to_double(const unsigned long long* __restrict__ w,
                                double* __restrict__ out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __longlong_as_double((long long)w[i]);
}

// ---- any nz: the inversion re-reads qh from global memory (cache hits)

// From qh (all layers) build ph, uh, vh of layer z at this thread's points and
// inverse-transform q, u, v along y into Qm, Um, Vm.
template <int NY, int E>
__device__ __forceinline__ void col_prologue_layer(const cd* __restrict__ qh,
                                                   cd* __restrict__ Qm, cd* __restrict__ Um,
                                                   cd* __restrict__ Vm, const MemberParams& P,
                                                   int nz, int z, int nk, int ldk, int t, int c, int k,
                                                   bool ok, size_t mbase, cd* lds,
                                                   const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  const size_t plane = (size_t)NY * ldk;
  const size_t zoff = mbase + (size_t)z * plane;
  const int kc = ok ? k : 0;
  cd x[G::E];
  cd ph[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const size_t pt = (size_t)(t + s * G::T) * ldk + kc;
    x[s] = ok ? qh[zoff + pt] : cd{0.0, 0.0};
    ph[s] = ok ? invert_point(P, nz, z, qh + mbase + pt, plane, pt) : cd{0.0, 0.0};
  }
  qgk::fft_block<NY, G::E, +1, G::TK>(x, lds, t, c, tw);
  {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Qm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
  }
  // uh = -(i ll) ph
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const double ll = P.ll[t + s * G::T];
    x[s] = {ll * ph[s].y, -ll * ph[s].x};
  }
  qgk::fft_block<NY, G::E, +1, G::TK>(x, lds, t, c, tw);
  {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Um + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
  }
  // vh = (i kk) ph
  const double kk = P.kk[kc];
#pragma unroll
  for (int s = 0; s < G::E; ++s) x[s] = {-kk * ph[s].y, kk * ph[s].x};
  qgk::fft_block<NY, G::E, +1, G::TK>(x, lds, t, c, tw);
  {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Vm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
  }
}

template <int NY, int E>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_prologue(
    const cd* __restrict__ qh, cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, int nz, int nk, int ldk, const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  __shared__ cd lds[NY * G::TK];
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int k = blockIdx.x * G::TK + c;
  const int m = blockIdx.y;
  const bool ok = k < nk;
  const MemberParams& P = params[m];
  const size_t mbase = (size_t)m * nz * NY * ldk;
  for (int z = 0; z < nz; ++z)
    col_prologue_layer<NY, E>(qh, Qm, Um, Vm, P, nz, z, nk, ldk, t, c, k, ok, mbase, lds, tw);
}

// Column phase of the fused step.  Per (member, column tile), layer by layer:
//  y-FFT of UQm, VQm -> S = ik uqh + il vqh; ph = a qh_old; dqhdt = -(S + ikQy ph)
//  [+ drag on the bottom layer]; AB3 + filter -> qh_new; the new dqhdt goes into
//  the dqhdt_pp buffer (it becomes dqhdt_p).  qh is double-buffered, so every
//  layer's inversion reads the old state while new layers are written.  Then
//  the next step's prologue (inversion of qh_new, y IFFT of q, u, v) runs.
template <int NY, int E>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_step(
    const cd* __restrict__ qh_old, cd* __restrict__ qh_new, const cd* __restrict__ p,
    cd* __restrict__ pp_to_newp, const cd* __restrict__ UQm, const cd* __restrict__ VQm,
    cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, StepCoeffs cf, int nz, int nk, int ldk,
    const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  __shared__ cd lds[NY * G::TK];
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int k = blockIdx.x * G::TK + c;
  const int m = blockIdx.y;
  const bool ok = k < nk;
  const int kc = ok ? k : 0;
  const MemberParams& P = params[m];
  const size_t plane = (size_t)NY * ldk;
  const size_t mbase = (size_t)m * nz * plane;
  const double kk = P.kk[kc];

  for (int z = 0; z < nz; ++z) {
    const size_t zoff = mbase + (size_t)z * plane;
    cd xa[G::E], xb[G::E];
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const size_t pt = (size_t)(t + s * G::T) * ldk + kc;
      xa[s] = ok ? ld_stream(UQm + zoff + pt) : cd{0.0, 0.0};
      xb[s] = ok ? ld_stream(VQm + zoff + pt) : cd{0.0, 0.0};
    }
    qgk::fft_block<NY, G::E, -1, G::TK>(xa, lds, t, c, tw);
    qgk::fft_block<NY, G::E, -1, G::TK>(xb, lds, t, c, tw);
    if (ok) {
      const double kq = kk * P.Qy[z];
      const bool drag = (z == nz - 1) && P.rek != 0.0;
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const int l = t + s * G::T;
        const size_t pt = (size_t)l * ldk + k;
        const size_t o = zoff + pt;
        const double ll = P.ll[l];
        // ik uqh + il vqh (kernel.pyx:391-392)
        const cd S = {-kk * xa[s].y + -ll * xb[s].y, kk * xa[s].x + ll * xb[s].x};
        const cd ph = invert_point(P, nz, z, qh_old + mbase + pt, plane, pt);
        cd d = {S.x + -kq * ph.y, S.y + kq * ph.x};
        d = {-d.x, -d.y};
        if (drag) {
          const double cr = P.rek * P.k2l2[pt];
          d = {d.x + cr * ph.x, d.y + cr * ph.y};
        }
        const double f = P.filtr[pt];
        const cd q = qh_old[o], pv = ld_stream(p + o), ppv = ld_stream(pp_to_newp + o);
        double rx = q.x + cf.dt1 * d.x;
        double ry = q.y + cf.dt1 * d.y;
        rx = rx + cf.dt2 * pv.x;
        ry = ry + cf.dt2 * pv.y;
        rx = rx + cf.dt3 * ppv.x;
        ry = ry + cf.dt3 * ppv.y;
        qh_new[o] = {f * rx, f * ry};
        st_stream(pp_to_newp + o, d);
      }
    } else {
      // lanes past nk write zeros so every tile row is a full 128-byte line
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const size_t o = zoff + (size_t)(t + s * G::T) * ldk + k;
        qh_new[o] = cd{0.0, 0.0};
        st_stream(pp_to_newp + o, cd{0.0, 0.0});
      }
    }
  }
  for (int z = 0; z < nz; ++z)
    col_prologue_layer<NY, E>(qh_new, Qm, Um, Vm, P, nz, z, nk, ldk, t, c, k, ok, mbase, lds, tw);
}

// ---- 3 <= nz <= 16: ph kept in its own buffer.  The prologue that produces
// the next step's Qm, Um, Vm also writes ph, so the tendency reads ph once; the
// inversion reads each layer's qh once per point (in registers) right after
// the block wrote it.  Every thread reads back only the points it wrote, so no
// barrier separates the phases.
constexpr int kTileNzMax = 16;

// phm = a qh at this thread's points of the column tile, for every layer, in
// invert_point's summation order.
template <int NY, int E>
__device__ __forceinline__ void col_invert_points(const cd* __restrict__ qh, cd* __restrict__ phm,
                                                  const MemberParams& P, int nz, int t, int k,
                                                  bool ok, size_t mbase, int ldk) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  if (!ok) return;
  const size_t plane = (size_t)NY * ldk;
#pragma unroll 1
  for (int s = 0; s < G::E; ++s) {
    const size_t pt = (size_t)(t + s * G::T) * ldk + k;
    if (P.tri) {
      // Elimination down the layers and back-substitution: S - kappa^2 I is
      // tridiagonal and, for kappa > 0, strictly diagonally dominant (S has a
      // negative diagonal and rows that sum to zero), so no pivoting is
      // needed.  ph = 0 at kappa = 0, where the reference zeroes a.
      const double w = gld(P.k2l2, pt);
      if (w == 0.0) {
        for (int z = 0; z < nz; ++z) phm[mbase + z * plane + pt] = cd{0.0, 0.0};
        continue;
      }
      const double* lo = P.tri;
      const double* dg = P.tri + nz;
      const double* hi = P.tri + 2 * nz;
      double cp[kTileNzMax];
      cd dp[kTileNzMax];
      double c_prev = 0.0;
      cd d_prev = {0.0, 0.0};
#pragma unroll
      for (int j = 0; j < kTileNzMax; ++j) {
        if (j < nz) {
          const cd qj = qh[mbase + j * plane + pt];
          const double l = gld(lo, j);
          const double r = 1.0 / ((gld(dg, j) - w) - l * c_prev);
          c_prev = gld(hi, j) * r;
          d_prev = {(qj.x - l * d_prev.x) * r, (qj.y - l * d_prev.y) * r};
          cp[j] = c_prev;
          dp[j] = d_prev;
        }
      }
      cd xn = {0.0, 0.0};
#pragma unroll
      for (int j = kTileNzMax - 1; j >= 0; --j) {
        if (j < nz) {
          xn = {dp[j].x - cp[j] * xn.x, dp[j].y - cp[j] * xn.y};
          phm[mbase + j * plane + pt] = xn;
        }
      }
      continue;
    }
    if (!P.a) {  // closed-form two-layer inverse (not used by layered buckets)
      for (int z = 0; z < nz; ++z)
        phm[mbase + z * plane + pt] = invert_point(P, nz, z, qh + mbase + pt, plane, pt);
      continue;
    }
    cd q[kTileNzMax];
#pragma unroll
    for (int j = 0; j < kTileNzMax; ++j)
      q[j] = j < nz ? qh[mbase + j * plane + pt] : cd{0.0, 0.0};
#pragma unroll
    for (int z = 0; z < kTileNzMax; ++z) {
      if (z < nz) {
        cd acc = {0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kTileNzMax; ++j) {
          if (j < nz) {
            const double a = gld(P.a, ((size_t)z * nz + j) * plane + pt);
            acc.x = acc.x + a * q[j].x;
            acc.y = acc.y + a * q[j].y;
          }
        }
        phm[mbase + z * plane + pt] = acc;
      }
    }
  }
}

// y IFFT of q, u = -il ph, v = ik ph of layer z from qh and the ph buffer
// (WQ false: u and v only, the step kernel having transformed q already)
template <int NY, int E, int TAG = 0, bool WQ = true>
__device__ __forceinline__ void col_prologue_layer_ph(const cd* __restrict__ qh,
                                                      const cd* __restrict__ phm,
                                                      cd* __restrict__ Qm, cd* __restrict__ Um,
                                                      cd* __restrict__ Vm, const MemberParams& P,
                                                      int z, int ldk, int t, int c, int k,
                                                      bool ok, size_t mbase, cd* lds,
                                                      const cd* __restrict__ tw) {
  using G = FGeo<NY, E>;
  const size_t plane = (size_t)NY * ldk;
  const size_t zoff = mbase + (size_t)z * plane;
  const int kc = ok ? k : 0;
  cd x[G::E], ph[G::E];
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const size_t pt = (size_t)(t + s * G::T) * ldk + kc;
    if constexpr (WQ) x[s] = ok ? qh[zoff + pt] : cd{0.0, 0.0};
    ph[s] = ok ? phm[zoff + pt] : cd{0.0, 0.0};
  }
  if constexpr (WQ) {
    qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
    {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
      for (int s = 0; s < G::E; ++s) st_stream(Qm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
    }
  }
#pragma unroll
  for (int s = 0; s < G::E; ++s) {
    const double ll = gld(P.ll, t + s * G::T);
    x[s] = {ll * ph[s].y, -ll * ph[s].x};
  }
  qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
  {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Um + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
  }
  const double kk = gld(P.kk, kc);
#pragma unroll
  for (int s = 0; s < G::E; ++s) x[s] = {-kk * ph[s].y, kk * ph[s].x};
  qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
  {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Vm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
  }
}

template <int NY, int E>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_prologue_ph(
    const cd* __restrict__ qh, cd* __restrict__ phm, cd* __restrict__ Qm, cd* __restrict__ Um,
    cd* __restrict__ Vm, const MemberParams* __restrict__ params, int nz, int nk, int ldk,
    const cd* __restrict__ tw) {
  using G = FGeo<NY, E>;
  __shared__ cd lds[NY * G::TK];
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int k = blockIdx.x * G::TK + c;
  const int m = blockIdx.y;
  const bool ok = k < nk;
  const MemberParams& P = params[m];
  const size_t mbase = (size_t)m * nz * NY * ldk;
  col_invert_points<NY, E>(qh, phm, P, nz, t, k, ok, mbase, ldk);
  for (int z = 0; z < nz; ++z)
    col_prologue_layer_ph<NY, E>(qh, phm, Qm, Um, Vm, P, z, ldk, t, c, k, ok, mbase, lds, tw);
}

// Column phase for 3 <= nz <= 16: as k_col_step, with ph_old read from the ph
// buffer, then ph_new = a qh_new written there for the next step.
template <int NY, int E>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_step_ph(
    const cd* __restrict__ qh_old, cd* __restrict__ qh_new, const cd* __restrict__ p,
    cd* __restrict__ pp_to_newp, const cd* __restrict__ UQm, const cd* __restrict__ VQm,
    cd* __restrict__ phm, cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, StepCoeffs cf, int nz, int nk, int ldk,
    const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  __shared__ cd lds[NY * G::TK];
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int k = blockIdx.x * G::TK + c;
  const int m = blockIdx.y;
  const bool ok = k < nk;
  const int kc = ok ? k : 0;
  const MemberParams& P = params[m];
  const size_t plane = (size_t)NY * ldk;
  const size_t mbase = (size_t)m * nz * plane;
  const double kk = gld(P.kk, kc);

  for (int z = 0; z < nz; ++z) {
    const size_t zoff = mbase + (size_t)z * plane;
    cd xa[G::E], xb[G::E], qn[G::E];
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const size_t pt = (size_t)(t + s * G::T) * ldk + kc;
      xa[s] = ok ? ld_stream(UQm + zoff + pt) : cd{0.0, 0.0};
      xb[s] = ok ? ld_stream(VQm + zoff + pt) : cd{0.0, 0.0};
    }
    qgk::fft_block<NY, G::E, -1, G::TK, kTagColPh>(xa, lds, t, c, tw);
    qgk::fft_block<NY, G::E, -1, G::TK, kTagColPh>(xb, lds, t, c, tw);
    if (ok) {
      const double kq = kk * gld(P.Qy, z);
      const bool drag = (z == nz - 1) && P.rek != 0.0;
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const int l = t + s * G::T;
        const size_t pt = (size_t)l * ldk + k;
        const size_t o = zoff + pt;
        const double ll = gld(P.ll, l);
        const cd S = {-kk * xa[s].y + -ll * xb[s].y, kk * xa[s].x + ll * xb[s].x};
        const cd ph = ld_stream(phm + o);
        cd d = {S.x + -kq * ph.y, S.y + kq * ph.x};
        d = {-d.x, -d.y};
        if (drag) {
          const double cr = P.rek * gld(P.k2l2, pt);
          d = {d.x + cr * ph.x, d.y + cr * ph.y};
        }
        const double f = gld(P.filtr, pt);
        const cd q = ld_stream(qh_old + o), pv = ld_stream(p + o), ppv = ld_stream(pp_to_newp + o);
        double rx = q.x + cf.dt1 * d.x;
        double ry = q.y + cf.dt1 * d.y;
        rx = rx + cf.dt2 * pv.x;
        ry = ry + cf.dt2 * pv.y;
        rx = rx + cf.dt3 * ppv.x;
        ry = ry + cf.dt3 * ppv.y;
        qn[s] = {f * rx, f * ry};
        qh_new[o] = qn[s];  // read back by the inversion below
        st_stream(pp_to_newp + o, d);
      }
    } else {
      // lanes past nk write zeros so every tile row is a full 128-byte line
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const size_t o = zoff + (size_t)(t + s * G::T) * ldk + k;
        qn[s] = cd{0.0, 0.0};
        qh_new[o] = cd{0.0, 0.0};
        st_stream(pp_to_newp + o, cd{0.0, 0.0});
      }
    }
    // the next step's Qm of this layer from the new qh still in registers
    // (the prologue below then transforms u and v only)
    qgk::fft_block<NY, G::E, +1, G::TK, kTagColPh>(qn, lds, t, c, tw);
#pragma unroll
    for (int s = 0; s < G::E; ++s) st_stream(Qm + zoff + (size_t)(t + s * G::T) * ldk + k, qn[s]);
  }
  col_invert_points<NY, E>(qh_new, phm, P, nz, t, k, ok, mbase, ldk);
  for (int z = 0; z < nz; ++z)
    col_prologue_layer_ph<NY, E, kTagColPh, false>(qh_new, phm, Qm, Um, Vm, P, z, ldk, t, c, k,
                                                   ok, mbase, lds, tw);
}

// ---- nz <= 2: every field is read exactly once; the layers' qh and ph at the
// thread's points stay in registers.


// ph[z] = sum_k1 a[z,k1] q[k1] with the reference's summation order
// KW: kappa^2 at the point is given (wk) instead of read from the k2l2 table
template <int NZ, bool KW = false>
__device__ __forceinline__ cd invert_regs(const MemberParams& P, int z, const cd* q,
                                          size_t plane, size_t pt, double wk = 0.0) {
#pragma clang fp contract(off)
  if (P.a) {
    cd acc = {0.0, 0.0};
#pragma unroll
    for (int k1 = 0; k1 < NZ; ++k1) {
      const double a = gld(P.a, ((size_t)z * NZ + k1) * plane + pt);
      acc.x = acc.x + a * q[k1].x;
      acc.y = acc.y + a * q[k1].y;
    }
    return acc;
  }
  if constexpr (NZ == 2) {
    const double w = KW ? wk : gld(P.k2l2, pt);
    const double det = w * (w + P.F1 + P.F2);
    if (det == 0.0) return {0.0, 0.0};
    const double di = 1.0 / det;
    const double a0 = z == 0 ? -(w + P.F2) * di : -P.F2 * di;
    const double a1 = z == 0 ? -P.F1 * di : -(w + P.F1) * di;
    cd acc;
    acc.x = 0.0 + a0 * q[0].x;
    acc.y = 0.0 + a0 * q[0].y;
    acc.x = acc.x + a1 * q[1].x;
    acc.y = acc.y + a1 * q[1].y;
    return acc;
  } else {
    return {0.0, 0.0};
  }
}

// Both layers' ph at one point (two layers), with invert_regs' arithmetic:
// the closed form's determinant and its reciprocal are formed once.
template <bool KW = false>
__device__ __forceinline__ void invert2_both(const MemberParams& P, cd q0, cd q1, size_t plane,
                                             size_t pt, cd& ph0, cd& ph1, double wk = 0.0) {
#pragma clang fp contract(off)
  if (P.a) {
    const cd q[2] = {q0, q1};
    ph0 = invert_regs<2>(P, 0, q, plane, pt);
    ph1 = invert_regs<2>(P, 1, q, plane, pt);
    return;
  }
  const double w = KW ? wk : gld(P.k2l2, pt);
  const double det = w * (w + P.F1 + P.F2);
  if (det == 0.0) {
    ph0 = {0.0, 0.0};
    ph1 = {0.0, 0.0};
    return;
  }
  const double di = 1.0 / det;
  const double a00 = -(w + P.F2) * di, a01 = -P.F1 * di;
  const double a10 = -P.F2 * di, a11 = -(w + P.F1) * di;
  ph0.x = 0.0 + a00 * q0.x;
  ph0.y = 0.0 + a00 * q0.y;
  ph0.x = ph0.x + a01 * q1.x;
  ph0.y = ph0.y + a01 * q1.y;
  ph1.x = 0.0 + a10 * q0.x;
  ph1.y = 0.0 + a10 * q0.y;
  ph1.x = ph1.x + a11 * q1.x;
  ph1.y = ph1.y + a11 * q1.y;
}

// inversion + y IFFT of q, u, v for all layers from register-held qh.
// PK (packed x-mixed rows): the lane of column 0 also carries the Nyquist
// column nk-1, whose spectra of q, u and v (all layers) it finds in
// nyq[(f * NZ + z) * NY + l] for f = 0, 1, 2 (every row, written before a
// barrier): it transforms column 0 + i * column nk-1 as one complex sequence.
// The c2r uses only the real parts of the y inverse transforms of these two
// columns, i.e. the transforms of their Hermitian parts H(l) = (X(l) +
// conj X(-l)) / 2, so those are what is packed.  Column 0 is Hermitian to
// rounding except in the self-conjugate rows 0 and NY/2, where the tendency's
// i l VQ term leaves an imaginary part (kept in the state, as the reference
// keeps it, and dropped by its c2r): there its real part is packed.  The
// Nyquist column's ik factor makes i k UQ and v anti-Hermitian in every row,
// so its Hermitian part is formed from the LDS copy.  The transform of the
// packed column is then real + i * real: the two x-mixed bins the row kernel
// takes.
template <int NY, int NZ, int E, int TAG = 0, bool PK = false>
__device__ __forceinline__ void col_prologue_regs(cd (&Q)[NZ][E],
                                                  cd* __restrict__ Qm, cd* __restrict__ Um,
                                                  cd* __restrict__ Vm, const MemberParams& P,
                                                  int nk, int ldk, int t, int c, int k, bool ok,
                                                  size_t mbase, cd* lds,
                                                  const cd* __restrict__ tw,
                                                  const cd* nyq = nullptr) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  const size_t plane = (size_t)NY * ldk;
  const int kc = ok ? k : 0;
  const double kk = gld(P.kk, kc);
  const bool dual = PK && ok && k == 0;
#pragma unroll
  for (int z = 0; z < NZ; ++z) {
    const size_t zoff = mbase + (size_t)z * plane;
    cd x[G::E], ph[G::E];
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      cd qs[NZ];
#pragma unroll
      for (int k1 = 0; k1 < NZ; ++k1) qs[k1] = Q[k1][s];
      const double ll = PK ? std_ll<NY>(P, t + s * G::T) : 0.0;
      ph[s] = ok ? invert_regs<NZ, PK>(P, z, qs, plane, (size_t)(t + s * G::T) * ldk + kc,
                                       kk * kk + ll * ll)
                 : cd{0.0, 0.0};
      x[s] = Q[z][s];
    }
    if constexpr (PK) {
      if (dual) {
#pragma unroll
        for (int s = 0; s < G::E; ++s) {
          const int l = t + s * G::T;
          const cd X = nyq[z * NY + l], Xm = nyq[z * NY + (l == 0 ? 0 : NY - l)];
          const cd b = {0.5 * (X.x + Xm.x), 0.5 * (X.y - Xm.y)};  // Hermitian part
          cd a = x[s];
          if (l == 0 || 2 * l == NY) a.y = 0.0;  // self-conjugate rows: real parts
          x[s] = {a.x - b.y, a.y + b.x};
        }
      }
    }
    qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
    {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
      for (int s = 0; s < G::E; ++s) st_stream(Qm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
    }
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const double ll = PK ? std_ll<NY>(P, t + s * G::T) : gld(P.ll, t + s * G::T);
      x[s] = {ll * ph[s].y, -ll * ph[s].x};
    }
    if constexpr (PK) {
      if (dual) {
#pragma unroll
        for (int s = 0; s < G::E; ++s) {
          const int l = t + s * G::T;
          const cd X = nyq[(NZ + z) * NY + l], Xm = nyq[(NZ + z) * NY + (l == 0 ? 0 : NY - l)];
          const cd b = {0.5 * (X.x + Xm.x), 0.5 * (X.y - Xm.y)};  // Hermitian part
          cd a = x[s];
          if (l == 0 || 2 * l == NY) a.y = 0.0;  // self-conjugate rows: real parts
          x[s] = {a.x - b.y, a.y + b.x};
        }
      }
    }
    qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
    {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
      for (int s = 0; s < G::E; ++s) st_stream(Um + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
    }
#pragma unroll
    for (int s = 0; s < G::E; ++s) x[s] = {-kk * ph[s].y, kk * ph[s].x};
    if constexpr (PK) {
      if (dual) {
#pragma unroll
        for (int s = 0; s < G::E; ++s) {
          const int l = t + s * G::T;
          const cd X = nyq[(2 * NZ + z) * NY + l], Xm = nyq[(2 * NZ + z) * NY + (l == 0 ? 0 : NY - l)];
          const cd b = {0.5 * (X.x + Xm.x), 0.5 * (X.y - Xm.y)};  // Hermitian part
          cd a = x[s];
          if (l == 0 || 2 * l == NY) a.y = 0.0;  // self-conjugate rows: real parts
          x[s] = {a.x - b.y, a.y + b.x};
        }
      }
    }
    qgk::fft_block<NY, G::E, +1, G::TK, TAG>(x, lds, t, c, tw);
    {  // whole tile rows (zeros past nk): full 128-byte lines
#pragma unroll
      for (int s = 0; s < G::E; ++s) st_stream(Vm + zoff + (size_t)(t + s * G::T) * ldk + k, x[s]);
    }
  }
}

// The layer-split kernel is used where its two transform buffers fit the
// 160 KB of LDS and it compiles without spills: its blocks are twice the
// register kernel's, so the register budget per thread halves (radix 4 fits
// at up to 1024 threads, radix 6 and 8 at up to 512, radix 12 and 16 at 256).
template <int NY, int E>
constexpr bool split_fits() {
  using G = FGeo<NY, E>;
  constexpr int threads = 2 * G::T * G::TK;
  constexpr bool regs = G::E <= 4 || (G::E <= 8 ? threads <= 512 : threads <= 256);
  return regs && 2 * NY * G::TK * (int)sizeof(cd) <= 160 * 1024;
}

// ---- nz = 2, one member: the layers split over the block.  A one-member
// launch is a handful of blocks whose time is the latency of their serial
// chain of transforms; here each layer has its own group of T x TK threads
// (the same column geometry and radix as k_col_step_reg), so the two layers'
// transforms run side by side instead of one after the other.  The inversion
// couples the layers at every point: the groups swap their qh through LDS
// (each group's transform buffer) before the tendency and again before the
// prologue.
template <int NY, int E>
__global__ __launch_bounds__((2 * FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_step_split(
    const cd* __restrict__ qh_old, cd* __restrict__ qh_new, const cd* __restrict__ p,
    cd* __restrict__ pp_to_newp, const cd* __restrict__ UQm, const cd* __restrict__ VQm,
    cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, StepCoeffs cf, int nk, int ldk,
    const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  constexpr int T = G::T, TK = G::TK, EE = G::E;
  __shared__ cd lds[2][NY * TK];
  const int lz = threadIdx.x / (T * TK);
  const int r = threadIdx.x % (T * TK);
  const int c = r % TK, t = r / TK;
  const int k = blockIdx.x * TK + c;
  const int m = blockIdx.y;
  const bool ok = k < nk;
  const int kc = ok ? k : 0;
  const MemberParams& P = params[m];
  const size_t plane = (size_t)NY * ldk;
  const size_t mbase = (size_t)m * 2 * plane;
  const size_t zoff = mbase + (size_t)lz * plane;
  const double kk = P.kk[kc];
  cd* L = lds[lz];
  const cd* Lo = lds[1 - lz];

  // own layer's old qh and the tendency's transforms of UQm, VQm
  cd Q[EE], xa[EE], xb[EE];
#pragma unroll
  for (int s = 0; s < EE; ++s) {
    const size_t pt = (size_t)(t + s * T) * ldk + kc;
    Q[s] = ok ? ld_stream(qh_old + zoff + pt) : cd{0.0, 0.0};
    xa[s] = ok ? ld_stream(UQm + zoff + pt) : cd{0.0, 0.0};
    xb[s] = ok ? ld_stream(VQm + zoff + pt) : cd{0.0, 0.0};
  }
  qgk::fft_block<NY, EE, -1, TK, kTagColSplit>(xa, L, t, c, tw);
  qgk::fft_block<NY, EE, -1, TK, kTagColSplit>(xb, L, t, c, tw);
  __syncthreads();  // the transforms' last LDS reads are done
#pragma unroll
  for (int s = 0; s < EE; ++s) L[(t + s * T) * TK + c] = Q[s];
  __syncthreads();
  // -(ik uqh + il vqh + ikQy ph) [+ drag], AB3 + filter
  if (ok) {
    const double kq = kk * P.Qy[lz];
    const bool drag = lz == 1 && P.rek != 0.0;
#pragma unroll
    for (int s = 0; s < EE; ++s) {
      const int l = t + s * T;
      const size_t pt = (size_t)l * ldk + k;
      const size_t o = zoff + pt;
      const cd qo = Lo[l * TK + c];
      const cd qs[2] = {lz == 0 ? Q[s] : qo, lz == 0 ? qo : Q[s]};
      const cd ph = invert_regs<2>(P, lz, qs, plane, pt);
      const double ll = P.ll[l];
      const cd S = {-kk * xa[s].y + -ll * xb[s].y, kk * xa[s].x + ll * xb[s].x};
      cd d = {S.x + -kq * ph.y, S.y + kq * ph.x};
      d = {-d.x, -d.y};
      if (drag) {
        const double cr = P.rek * P.k2l2[pt];
        d = {d.x + cr * ph.x, d.y + cr * ph.y};
      }
      const double f = P.filtr[pt];
      const cd q = Q[s], pv = ld_stream(p + o), ppv = ld_stream(pp_to_newp + o);
      double rx = q.x + cf.dt1 * d.x;
      double ry = q.y + cf.dt1 * d.y;
      rx = rx + cf.dt2 * pv.x;
      ry = ry + cf.dt2 * pv.y;
      rx = rx + cf.dt3 * ppv.x;
      ry = ry + cf.dt3 * ppv.y;
      const cd qn = {f * rx, f * ry};
      st_stream(qh_new + o, qn);
      st_stream(pp_to_newp + o, d);
      Q[s] = qn;
    }
  } else {
    // lanes past nk write zeros so every tile row is a full 128-byte line
#pragma unroll
    for (int s = 0; s < EE; ++s) {
      const size_t o = zoff + (size_t)(t + s * T) * ldk + k;
      st_stream(qh_new + o, cd{0.0, 0.0});
      st_stream(pp_to_newp + o, cd{0.0, 0.0});
    }
  }
  // next step's prologue: ph from both layers' new qh, then y IFFT of q, u, v
  __syncthreads();  // every read of the old qh swap is done
#pragma unroll
  for (int s = 0; s < EE; ++s) L[(t + s * T) * TK + c] = Q[s];
  __syncthreads();
  cd ph[EE];
#pragma unroll
  for (int s = 0; s < EE; ++s) {
    const cd qo = Lo[(t + s * T) * TK + c];
    const cd qs[2] = {lz == 0 ? Q[s] : qo, lz == 0 ? qo : Q[s]};
    ph[s] = ok ? invert_regs<2>(P, lz, qs, plane, (size_t)(t + s * T) * ldk + kc)
               : cd{0.0, 0.0};
  }
  __syncthreads();  // every read of the new qh swap is done before L is reused
  cd x[EE];
#pragma unroll
  for (int s = 0; s < EE; ++s) x[s] = Q[s];
  qgk::fft_block<NY, EE, +1, TK, kTagColSplit>(x, L, t, c, tw);
#pragma unroll
  for (int s = 0; s < EE; ++s) st_stream(Qm + zoff + (size_t)(t + s * T) * ldk + k, x[s]);
#pragma unroll
  for (int s = 0; s < EE; ++s) {
    const double ll = P.ll[t + s * T];
    x[s] = {ll * ph[s].y, -ll * ph[s].x};
  }
  qgk::fft_block<NY, EE, +1, TK, kTagColSplit>(x, L, t, c, tw);
#pragma unroll
  for (int s = 0; s < EE; ++s) st_stream(Um + zoff + (size_t)(t + s * T) * ldk + k, x[s]);
#pragma unroll
  for (int s = 0; s < EE; ++s) x[s] = {-kk * ph[s].y, kk * ph[s].x};
  qgk::fft_block<NY, EE, +1, TK, kTagColSplit>(x, L, t, c, tw);
#pragma unroll
  for (int s = 0; s < EE; ++s) st_stream(Vm + zoff + (size_t)(t + s * T) * ldk + k, x[s]);
}

// ll[l] on a standard grid (the packed register kernels): dl * l for
// l < NY/2, dl * (l - NY) above, as the wavenumber table is built
// (model.py:461-463), so the values are the table's to the bit.
template <int NY>
__device__ __forceinline__ double std_ll(const MemberParams& P, int l) {
  return P.dl * (double)(l < NY / 2 ? l : l - NY);
}

// Tendency and AB3 update at one spectral point (kernel.pyx:385-394, 444-461,
// 496-517): S = ik UQ + il VQ from the y-transformed fluxes, dqhdt = -(S + ik Qy
// ph) (+ rek kappa^2 ph in the bottom layer), qh_new = filtr (qh + dt1 dqhdt +
// dt2 dqhdt_p + dt3 dqhdt_pp), in the reference's operation order.
// With f_pt, the q parameterization's dqh at the point is added after the
// drag (kernel.pyx:424-442, called after _do_friction by model.py:420-426).
template <bool KW = false>
__device__ __forceinline__ cd tend_ab3_point(const MemberParams& P, cd UQ, cd VQ, cd ph, cd q,
                                             double kk, double kq, double ll, bool drag,
                                             size_t pt, const cd* __restrict__ p_pt,
                                             const cd* __restrict__ pp_pt, StepCoeffs cf,
                                             cd& d_out, double wk = 0.0,
                                             const cd* __restrict__ f_pt = nullptr) {
#pragma clang fp contract(off)
  const cd S = {-kk * UQ.y + -ll * VQ.y, kk * UQ.x + ll * VQ.x};
  cd d = {S.x + -kq * ph.y, S.y + kq * ph.x};
  d = {-d.x, -d.y};
  if (drag) {
    const double cr = P.rek * (KW ? wk : gld(P.k2l2, pt));
    d = {d.x + cr * ph.x, d.y + cr * ph.y};
  }
  if (f_pt) {
    const cd f = ld_stream(f_pt);
    d = {d.x + f.x, d.y + f.y};
  }
  const double f = gld(P.filtr, pt);
  const cd pv = ld_stream(p_pt), ppv = ld_stream(pp_pt);
  double rx = q.x + cf.dt1 * d.x;
  double ry = q.y + cf.dt1 * d.y;
  rx = rx + cf.dt2 * pv.x;
  ry = ry + cf.dt2 * pv.y;
  rx = rx + cf.dt3 * ppv.x;
  ry = ry + cf.dt3 * ppv.y;
  d_out = d;
  return {f * rx, f * ry};
}

// Columns handled by a launch of the register column kernels: nk, or nk - 1
// with packed x-mixed rows (the Nyquist column rides on column 0).
__host__ __device__ constexpr int col_count(int nk, bool pk) { return pk ? nk - 1 : nk; }

// The Nyquist column (nk-1) after the update of every layer: its new qh
// (nyq[0][z][l], written by the step or loaded by the prologue kernel) gives
// ph and so the y-spectra of u and v (nyq[1][z][l], nyq[2][z][l]), which the
// packed prologue adds to column 0's as imaginary parts.  All threads of the
// block take rows l = threadIdx.x, threadIdx.x + blockDim.x, ... (the same
// rows as the step's Nyquist update, so a thread reads only what it wrote).
template <int NY, int NZ>
__device__ __forceinline__ void nyquist_uv(const MemberParams& P, cd* nyq, int nk, int ldk) {
#pragma clang fp contract(off)
  const size_t plane = (size_t)NY * ldk;
  const double kkN = gld(P.kk, nk - 1);
  for (int l = threadIdx.x; l < NY; l += blockDim.x) {
    const size_t ptN = (size_t)l * ldk + (nk - 1);
    const double ll = std_ll<NY>(P, l);
    const double w = kkN * kkN + ll * ll;
    cd qs[NZ];
#pragma unroll
    for (int k1 = 0; k1 < NZ; ++k1) qs[k1] = nyq[k1 * NY + l];
#pragma unroll
    for (int z = 0; z < NZ; ++z) {
      const cd ph = invert_regs<NZ, true>(P, z, qs, plane, ptN, w);
      nyq[(NZ + z) * NY + l] = {ll * ph.y, -ll * ph.x};
      nyq[(2 * NZ + z) * NY + l] = {-kkN * ph.y, kkN * ph.x};
    }
  }
}

template <int NY, int NZ, int E, bool PK>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK)) void k_col_prologue_reg(
    const cd* __restrict__ qh, cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, int nk, int ldk, const cd* __restrict__ tw) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  __shared__ cd lds[NY * G::TK];
  __shared__ cd nyq[PK ? 3 * NZ * NY : 1];
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int m = blockIdx.y;
  const int tile = PK ? (blockIdx.x + m) % gridDim.x : blockIdx.x;  // as k_col_step_reg
  const int k = tile * G::TK + c;
  const bool ok = k < col_count(nk, PK);
  const int kc = ok ? k : 0;
  const MemberParams& P = params[m];
  const size_t plane = (size_t)NY * ldk;
  const size_t mbase = (size_t)m * NZ * plane;
  cd Q[NZ][G::E];
#pragma unroll
  for (int z = 0; z < NZ; ++z)
#pragma unroll
    for (int s = 0; s < G::E; ++s)
      Q[z][s] = ok ? ld_stream(qh + mbase + z * plane + (size_t)(t + s * G::T) * ldk + kc)
                   : cd{0.0, 0.0};
  if constexpr (PK) {
    if (tile == 0) {  // the Nyquist column, for the lane that packs it
      for (int l = threadIdx.x; l < NY; l += blockDim.x)
#pragma unroll
        for (int z = 0; z < NZ; ++z)
          nyq[z * NY + l] = ld_stream(qh + mbase + z * plane + (size_t)l * ldk + (nk - 1));
      nyquist_uv<NY, NZ>(P, nyq, nk, ldk);
      __syncthreads();  // every row of nyq is written
    }
  }
  col_prologue_regs<NY, NZ, G::E, 0, PK>(Q, Qm, Um, Vm, P, nk, ldk, t, c, k, ok, mbase, lds, tw,
                                          nyq);
}

// Waves per SIMD the register column kernel is compiled for: with two layers
// at radix 8 the compiler would otherwise take all 256 VGPRs (one wave per
// SIMD); bounded to two waves it fits without spills and runs 10 % faster.
template <int NZ, int E>
constexpr int col_reg_waves() { return NZ == 2 && E == 8 ? 2 : 1; }

// PK: packed x-mixed rows (see k_row_flux).  The launch covers columns
// 0..nk-2; the lane of column 0 transforms the packed slot, column 0 +
// i * column nk-1, and separates the two spectra by their symmetry:
// F = A + iB with A, B Hermitian in l gives A(l) = (F(l) + conj F(-l)) / 2,
// B(l) = (F(l) - conj F(-l)) / 2i.  F is exchanged through LDS columns 0-1 of
// the transform buffer, A and B go to columns 2-5.  After the block's own
// columns, every thread of block 0 steps rows of the Nyquist column (B from
// LDS; its qh, dqhdt_p and dqhdt_pp from memory) and keeps the new qh and the
// derived u, v spectra in LDS for the packed prologue.  The next transform's
// first LDS write follows a barrier, so columns 2-5 live until then.
// FORCE: dqh (the q parameterization's spectral forcing, [M][nz][ny][ldk])
// is added to the tendency after the drag.
template <int NY, int NZ, int E, bool PK, bool FORCE = false>
__global__ __launch_bounds__((FGeo<NY, E>::T * FGeo<NY, E>::TK), (col_reg_waves<NZ, E>())) void k_col_step_reg(
    const cd* __restrict__ qh_old, cd* __restrict__ qh_new, const cd* __restrict__ p,
    cd* __restrict__ pp_to_newp, const cd* __restrict__ UQm, const cd* __restrict__ VQm,
    cd* __restrict__ Qm, cd* __restrict__ Um, cd* __restrict__ Vm,
    const MemberParams* __restrict__ params, StepCoeffs cf, int nk, int ldk,
    const cd* __restrict__ tw, const cd* __restrict__ dqh = nullptr) {
#pragma clang fp contract(off)
  using G = FGeo<NY, E>;
  static_assert(!PK || G::TK >= 6, "the packed exchange uses six LDS columns");
  __shared__ cd lds[NY * G::TK];
  __shared__ cd twl[NY];
  __shared__ cd nyq[PK ? 3 * NZ * NY : 1];
  tw = stage_twiddles<NY>(twl, tw);
  const int c = threadIdx.x % G::TK;
  const int t = threadIdx.x / G::TK;
  const int m = blockIdx.y;
  // packed: tiles rotated by member, so the tiles carrying the Nyquist column
  // (tile 0) spread over the XCDs (blocks go to XCDs round robin, and with 16
  // tiles per member an unrotated tile 0 would always land on the same one)
  const int tile = PK ? (blockIdx.x + m) % gridDim.x : blockIdx.x;
  const int k = tile * G::TK + c;
  const bool ok = k < col_count(nk, PK);
  const int kc = ok ? k : 0;
  const MemberParams& P = params[m];
  const size_t plane = (size_t)NY * ldk;
  const size_t mbase = (size_t)m * NZ * plane;
  const double kk = gld(P.kk, kc);
  const bool dual = PK && ok && k == 0;
  const int kN = nk - 1;

  cd Q[NZ][G::E];
  // two layers: layer 1's ph from the old qh of both layers, formed together
  // with layer 0's (one determinant per point) before layer 0 is replaced
  cd ph1[NZ == 2 ? G::E : 1];
#pragma unroll
  for (int z = 0; z < NZ; ++z)
#pragma unroll
    for (int s = 0; s < G::E; ++s)
      Q[z][s] = ok ? ld_stream(qh_old + mbase + z * plane + (size_t)(t + s * G::T) * ldk + kc)
                   : cd{0.0, 0.0};

#pragma unroll
  for (int z = 0; z < NZ; ++z) {
    const size_t zoff = mbase + (size_t)z * plane;
    cd xa[G::E], xb[G::E];
#pragma unroll
    for (int s = 0; s < G::E; ++s) {
      const size_t pt = (size_t)(t + s * G::T) * ldk + kc;
      xa[s] = ok ? ld_stream(UQm + zoff + pt) : cd{0.0, 0.0};
      xb[s] = ok ? ld_stream(VQm + zoff + pt) : cd{0.0, 0.0};
    }
    qgk::fft_block<NY, G::E, -1, G::TK, kTagColReg>(xa, lds, t, c, tw);
    qgk::fft_block<NY, G::E, -1, G::TK, kTagColReg>(xb, lds, t, c, tw);
    if constexpr (PK) {
      if (tile == 0) {  // block-uniform: the barriers below are safe
        __syncthreads();      // the transforms' last LDS reads are done
        if (dual) {
#pragma unroll
          for (int s = 0; s < G::E; ++s) {
            lds[t + s * G::T] = xa[s];
            lds[NY + t + s * G::T] = xb[s];
          }
        }
        __syncthreads();
        if (dual) {
#pragma nounroll
          for (int s = 0; s < G::E; ++s) {
            const int l = t + s * G::T;
            const int lm = l == 0 ? 0 : NY - l;  // -l mod NY
#pragma unroll
            for (int f = 0; f < 2; ++f) {
              const cd F = lds[f * NY + l], Fm = lds[f * NY + lm];
              lds[(2 + f) * NY + l] = {0.5 * (F.x + Fm.x), 0.5 * (F.y - Fm.y)};   // A
              lds[(4 + f) * NY + l] = {0.5 * (F.y + Fm.y), -0.5 * (F.x - Fm.x)};  // B
            }
          }
#pragma unroll
          for (int s = 0; s < G::E; ++s) {
            xa[s] = lds[2 * NY + t + s * G::T];
            xb[s] = lds[3 * NY + t + s * G::T];
          }
        }
      }
    }
    if (ok) {
      const double kq = kk * gld(P.Qy, z);
      const bool drag = (z == NZ - 1) && P.rek != 0.0;
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const int l = t + s * G::T;
        const size_t pt = (size_t)l * ldk + k;
        const size_t o = zoff + pt;
        // packed: a standard grid, ll and kappa^2 formed here (std_ll)
        const double ll = PK ? std_ll<NY>(P, l) : gld(P.ll, l);
        const double w = PK ? kk * kk + ll * ll : 0.0;
        // ph from the old qh of every layer
        cd ph;
        if constexpr (NZ == 2) {
          if (z == 0)
            invert2_both<PK>(P, Q[0][s], Q[1][s], plane, pt, ph, ph1[s], w);
          else
            ph = ph1[s];
        } else {
          ph = invert_regs<NZ, PK>(P, z, Q[0] + s, plane, pt, w);
        }
        cd d;
        const cd qn = tend_ab3_point<PK>(P, xa[s], xb[s], ph, Q[z][s], kk, kq, ll, drag, pt, p + o,
                                         pp_to_newp + o, cf, d, w, FORCE ? dqh + o : nullptr);
        st_stream(qh_new + o, qn);
        st_stream(pp_to_newp + o, d);
        Q[z][s] = qn;
      }
    } else if (!PK || k != kN) {
      // lanes past nk write zeros so every tile row is a full 128-byte line
      // (not over the Nyquist column, which the packed lane of column 0 steps)
#pragma unroll
      for (int s = 0; s < G::E; ++s) {
        const size_t o = zoff + (size_t)(t + s * G::T) * ldk + k;
        st_stream(qh_new + o, cd{0.0, 0.0});
        st_stream(pp_to_newp + o, cd{0.0, 0.0});
      }
    }
    if constexpr (PK) {
      if (tile == 0) {  // the Nyquist column's tendency and AB3 update
        __syncthreads();      // B of every row is in LDS
        const double kkN = gld(P.kk, kN);
        const double kqN = kkN * gld(P.Qy, z);
        const bool drag = (z == NZ - 1) && P.rek != 0.0;
        for (int l = threadIdx.x; l < NY; l += blockDim.x) {
          const size_t ptN = (size_t)l * ldk + kN;
          const double ll = std_ll<NY>(P, l);
          const double w = kkN * kkN + ll * ll;
          cd qs[NZ];
#pragma unroll
          for (int k1 = 0; k1 < NZ; ++k1) qs[k1] = ld_stream(qh_old + mbase + k1 * plane + ptN);
          const cd phN = invert_regs<NZ, true>(P, z, qs, plane, ptN, w);
          cd dN;
          const cd qnN = tend_ab3_point<true>(P, lds[4 * NY + l], lds[5 * NY + l], phN, qs[z], kkN,
                                              kqN, ll, drag, ptN, p + zoff + ptN,
                                              pp_to_newp + zoff + ptN, cf, dN, w,
                                              FORCE ? dqh + zoff + ptN : nullptr);
          st_stream(qh_new + zoff + ptN, qnN);
          st_stream(pp_to_newp + zoff + ptN, dN);
          nyq[z * NY + l] = qnN;
        }
      }
    }
  }
  if constexpr (PK) {
    if (tile == 0) {
      nyquist_uv<NY, NZ>(P, nyq, nk, ldk);
      __syncthreads();  // every row of nyq is written
    }
  }
  col_prologue_regs<NY, NZ, G::E, kTagColReg, PK>(Q, Qm, Um, Vm, P, nk, ldk, t, c, k, ok, mbase,
                                                   lds, tw, nyq);
}

// ==========================================================================
// Learned-closure support: coarse-graining between grids, NHWC packing
// ==========================================================================

// lo[m][z][l][k] = s * G[l][k] * hi[m][z][l'][k] for the modes both grids
// resolve (|l| < nlo/2, k < nlo/2); the coarse grid's Nyquist row and column
// are set to zero.  s = (nlo/nhi)^2 keeps physical amplitudes (the transforms
// are unnormalized sums over the grid).
__global__ void k_coarsen(const cd* __restrict__ hi, cd* __restrict__ lo,
                          const double* __restrict__ G, int nplanes, int nhi, int nlo,
                          double s) {
  const int nkh = ldk_of(nhi), nkl = ldk_of(nlo);
  const size_t total = (size_t)nplanes * nlo * nkl;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % nkl);
    const int l = (int)((i / nkl) % nlo);
    const size_t pl = i / ((size_t)nkl * nlo);
    cd v = {0.0, 0.0};
    if (k < nlo / 2 && l != nlo / 2) {
      const int lw = l < nlo / 2 ? l : l - nlo;
      const int lh = lw >= 0 ? lw : lw + nhi;
      const cd h = hi[(pl * nhi + lh) * nkh + k];
      const double g = G ? G[(size_t)l * nkl + k] * s : s;
      v = {h.x * g, h.y * g};
    }
    lo[i] = v;
  }
}

// a - b (spectral), into out
__global__ void k_csub(const cd* __restrict__ a, const cd* __restrict__ b, cd* __restrict__ out,
                       size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = {a[i].x - b[i].x, a[i].y - b[i].y};
}

// NHWC BF16 network input from real fields: channel c = f * nz + z of field f
// (f over q, u, v) scaled by inv_scale[c]; channels cin..cinp-1 are zero.
// One thread per pixel and group of 8 channels: the group is rounded to BF16
// (nearest even) and written as one 16-byte store; consecutive threads take
// consecutive pixels, so the field reads are coalesced.
__global__ void k_pack_nhwc(const double* __restrict__ q, const double* __restrict__ u,
                            const double* __restrict__ v, const double* __restrict__ inv_scale,
                            unsigned short* __restrict__ out, int M, int nz, int npix, int cinp) {
  const int groups = cinp / 8;
  const size_t total = (size_t)M * npix * groups;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t mp = i % ((size_t)M * npix);  // member * npix + pixel
    const int g = (int)(i / ((size_t)M * npix));
    const int p = (int)(mp % npix);
    const int m = (int)(mp / npix);
    unsigned short h[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int c = 8 * g + j;
      float val = 0.f;
      if (c < 3 * nz) {
        const int f = c / nz, z = c % nz;
        const double* src = f == 0 ? q : (f == 1 ? u : v);
        val = (float)(src[((size_t)m * nz + z) * npix + p] * inv_scale[c]);
      }
      unsigned int bits = __float_as_uint(val);
      bits += 0x7fffu + ((bits >> 16) & 1u);
      h[j] = (unsigned short)(bits >> 16);
    }
    uint4 w;
    w.x = h[0] | ((unsigned int)h[1] << 16);
    w.y = h[2] | ((unsigned int)h[3] << 16);
    w.z = h[4] | ((unsigned int)h[5] << 16);
    w.w = h[6] | ((unsigned int)h[7] << 16);
    *reinterpret_cast<uint4*>(out + mp * cinp + 8 * g) = w;
  }
}

// dq[m][z][p] = net[m][p][z] * scale[z] (network output, FP32 NHWC, coutp channels)
__global__ void k_unpack_forcing(const float* __restrict__ net, const double* __restrict__ scale,
                                 double* __restrict__ dq, int M, int nz, int npix, int coutp) {
  const size_t total = (size_t)M * nz * npix;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(i % npix);
    const int z = (int)((i / npix) % nz);
    const int m = (int)(i / ((size_t)npix * nz));
    dq[i] = (double)net[((size_t)m * npix + p) * coutp + z] * scale[z];
  }
}

// ==========================================================================
// Diagnostics (pyqg/model.py:644-736, qg_diagnostics.py:6-70,
// qg_model.py:213-263, layered_model.py:242-301)
// ==========================================================================
constexpr int kDiagNzSmall = 16;  // per-point arrays of the common case
constexpr int kDiagNzMax = 64;

// Offsets into the per-member diagnostic parameter block (doubles):
// Hi[nz], H, S[nz][nz], del1, del2, rd^-2, U1-U2, Vbg[nz], Qx[nz],
// f2gpi[nz] (nz-1 used), inverse of the vertical-mode matrix [nz][nz], kdi2[nz].
struct DiagLayout {
  int nz;
  __host__ __device__ int Hi() const { return 0; }
  __host__ __device__ int H() const { return nz; }
  __host__ __device__ int S() const { return nz + 1; }
  __host__ __device__ int del1() const { return nz + 1 + nz * nz; }
  __host__ __device__ int del2() const { return del1() + 1; }
  __host__ __device__ int rd2inv() const { return del1() + 2; }
  __host__ __device__ int Ushear() const { return del1() + 3; }
  __host__ __device__ int Vbg() const { return del1() + 4; }
  __host__ __device__ int Qx() const { return Vbg() + nz; }
  __host__ __device__ int f2gpi() const { return Qx() + nz; }
  __host__ __device__ int pinv() const { return f2gpi() + nz; }
  __host__ __device__ int kdi2() const { return pinv() + nz * nz; }
  __host__ __device__ int len() const { return kdi2() + nz; }
};

struct DiagArgs {
  const cd *qh, *ph, *D, *P, *PP;  // state [M][nz][nl][ldk]
  const cd *dqh, *duh, *dvh;       // parameterization tendencies (or null)
  const cd *Jq, *Jpxi, *J3;        // Jacobians; J3: Jptpc [M][1] (nz = 2) or JSp [M][nz]
  const double* prm;               // [M][plen]
  double* acc;                     // [slot][M][nl][ldk]
  double* pts;                     // per-point APEgen terms [M][nl][ldk] (or null)
  int M, nz, nl, nk, ldk, plen;
  double invM2, dt, dt1, dt2, dt3;
  int slot[QGK_NDIAG];             // first accumulator slot, -1 = not requested
};

// (a . v)_i at one spectral point: the member's inversion table or the
// closed-form two-layer inverse (qg_model.py:146-161)
__device__ cd inv_apply(const MemberParams& P, int nz, int i, const cd* v, size_t plane,
                        size_t pt) {
  if (P.a) {
    cd acc = {0.0, 0.0};
    for (int j = 0; j < nz; ++j) {
      const double a = P.a[((size_t)i * nz + j) * plane + pt];
      acc.x += a * v[j].x;
      acc.y += a * v[j].y;
    }
    return acc;
  }
  const double w = P.k2l2[pt];
  const double det = w * (w + P.F1 + P.F2);
  if (det == 0.0) return {0.0, 0.0};
  const double a0 = i == 0 ? -(w + P.F2) / det : -P.F2 / det;
  const double a1 = i == 0 ? -P.F1 / det : -(w + P.F1) / det;
  return {a0 * v[0].x + a1 * v[1].x, a0 * v[0].y + a1 * v[1].y};
}

__device__ __forceinline__ double re_conj_mul(cd a, cd b) {  // Re(conj(a) * b)
  return a.x * b.x + a.y * b.y;
}
__device__ __forceinline__ double abs2(cd a) { return a.x * a.x + a.y * a.y; }

// Every requested spectral diagnostic at every (member, l, k), added into its
// accumulator; the per-point APEgen terms go to pts for a per-member sum.
template <int NZMAX>
__global__ void k_diag_spectral(DiagArgs A, const MemberParams* __restrict__ mp) {
  const size_t plane = (size_t)A.nl * A.ldk;
  const size_t total = (size_t)A.M * plane;
  const int nz = A.nz;
  const DiagLayout L{nz};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / plane);
    const size_t pt = i % plane;
    const int l = (int)(pt / A.ldk), k = (int)(pt % A.ldk);
    if (k >= A.nk) {
      if (A.pts) A.pts[i] = 0.0;
      continue;
    }
    const MemberParams& P = mp[m];
    const double* pr = A.prm + (size_t)m * A.plen;
    const size_t base = (size_t)m * nz * plane + pt;
    cd qh[NZMAX], ph[NZMAX];
    for (int z = 0; z < nz; ++z) {
      qh[z] = A.qh[base + z * plane];
      ph[z] = A.ph[base + z * plane];
    }
    const double kk = P.kk[k], ll = P.ll[l], wv2 = P.k2l2[pt];
    const double H = pr[L.H()];
    const double* Hi = pr + L.Hi();
    const double s = A.invM2;
    auto add = [&](int d, int j, double v) {
      A.acc[((size_t)(A.slot[d] + j) * A.M + m) * plane + pt] += v;
    };
    const int bot = nz - 1;
    if (A.slot[QGK_D_ENSSPEC] >= 0)
      for (int z = 0; z < nz; ++z) add(QGK_D_ENSSPEC, z, abs2(qh[z]) * s);
    if (A.slot[QGK_D_KESPEC] >= 0)
      for (int z = 0; z < nz; ++z) add(QGK_D_KESPEC, z, wv2 * abs2(ph[z]) * s);
    if (A.slot[QGK_D_KEFRICTIONSPEC] >= 0)
      add(QGK_D_KEFRICTIONSPEC, 0, -P.rek * Hi[bot] / H * wv2 * abs2(ph[bot]) * s);
    if (A.slot[QGK_D_DISSSPEC] >= 0 || A.slot[QGK_D_ENSDISSSPEC] >= 0) {
      const double f1 = P.filtr[pt] - 1.0;
      double de = 0.0, dz = 0.0;
      for (int z = 0; z < nz; ++z) {
        const size_t o = base + z * plane;
        const cd d = A.D[o], p = A.P[o], pp = A.PP[o];
        cd r = {qh[z].x + A.dt1 * d.x + A.dt2 * p.x + A.dt3 * pp.x,
                qh[z].y + A.dt1 * d.y + A.dt2 * p.y + A.dt3 * pp.y};
        r = {f1 * r.x, f1 * r.y};
        de += Hi[z] * re_conj_mul(ph[z], r);
        dz += Hi[z] * re_conj_mul(qh[z], r);
      }
      if (A.slot[QGK_D_DISSSPEC] >= 0) add(QGK_D_DISSSPEC, 0, -de / H / A.dt * s);
      if (A.slot[QGK_D_ENSDISSSPEC] >= 0) add(QGK_D_ENSDISSSPEC, 0, dz / H / A.dt * s);
    }
    const bool want_param = A.slot[QGK_D_PARAMSPEC] >= 0 || A.slot[QGK_D_ENSPARAMSPEC] >= 0 ||
                            A.slot[QGK_D_PARAMSPEC_APEFLUX] >= 0 ||
                            A.slot[QGK_D_PARAMSPEC_KEFLUX] >= 0;
    if (want_param) {
      cd dp[NZMAX];
      for (int z = 0; z < nz; ++z) {
        const size_t o = base + z * plane;
        cd v = {0.0, 0.0};
        if (A.duh) {  // -il duh + ik dvh
          const cd du = A.duh[o], dv = A.dvh[o];
          v.x += ll * du.y - kk * dv.y;
          v.y += -ll * du.x + kk * dv.x;
        }
        if (A.dqh) {
          const cd dq = A.dqh[o];
          v.x += dq.x;
          v.y += dq.y;
        }
        dp[z] = v;
      }
      double ps = 0.0, es = 0.0, ape = 0.0, ke = 0.0;
      cd adp[NZMAX];
      const bool want_a = A.slot[QGK_D_PARAMSPEC_APEFLUX] >= 0 ||
                          A.slot[QGK_D_PARAMSPEC_KEFLUX] >= 0;
      if (want_a)
        for (int z = 0; z < nz; ++z) adp[z] = inv_apply(P, nz, z, dp, plane, pt);
      for (int z = 0; z < nz; ++z) {
        const double hr = Hi[z] / H;
        ps += hr * re_conj_mul(ph[z], dp[z]);
        es += hr * re_conj_mul(qh[z], dp[z]);
        if (want_a) {
          cd t = {0.0, 0.0};
          for (int j = 0; j < nz; ++j) {
            const double Sij = pr[L.S() + z * nz + j];
            t.x += Sij * adp[j].x;
            t.y += Sij * adp[j].y;
          }
          ape += hr * re_conj_mul(ph[z], t);
          ke += hr * re_conj_mul(ph[z], adp[z]);
        }
      }
      if (A.slot[QGK_D_PARAMSPEC] >= 0) add(QGK_D_PARAMSPEC, 0, -ps * s);
      if (A.slot[QGK_D_ENSPARAMSPEC] >= 0) add(QGK_D_ENSPARAMSPEC, 0, es * s);
      if (A.slot[QGK_D_PARAMSPEC_APEFLUX] >= 0) add(QGK_D_PARAMSPEC_APEFLUX, 0, -ape * s);
      if (A.slot[QGK_D_PARAMSPEC_KEFLUX] >= 0) add(QGK_D_PARAMSPEC_KEFLUX, 0, wv2 * ke * s);
    }
    if (A.slot[QGK_D_ENTSPEC] >= 0) {
      cd b = {0.0, 0.0};
      for (int z = 0; z < nz; ++z) {
        b.x += Hi[z] * qh[z].x;
        b.y += Hi[z] * qh[z].y;
      }
      b = {b.x / H, b.y / H};
      add(QGK_D_ENTSPEC, 0, abs2(b) * s);
    }
    if (A.slot[QGK_D_ENSFLUX] >= 0) {
      double e = 0.0;
      for (int z = 0; z < nz; ++z) e += -Hi[z] * re_conj_mul(qh[z], A.Jq[base + z * plane]);
      add(QGK_D_ENSFLUX, 0, e / H * s);
    }
    if (A.slot[QGK_D_ENSGENSPEC] >= 0 || A.slot[QGK_D_APEGENSPEC] >= 0) {
      double eg = 0.0, ag = 0.0;
      for (int z = 0; z < nz; ++z) {
        cd sph = {0.0, 0.0};
        for (int j = 0; j < nz; ++j) {
          const double Sij = pr[L.S() + z * nz + j];
          sph.x += Sij * ph[j].x;
          sph.y += Sij * ph[j].y;
        }
        // (ilQx - ikQy) conj(Sph) ph: a purely imaginary factor i*c
        const double c = ll * pr[L.Qx() + z] - kk * P.Qy[z];
        const cd w = cmul(cconj(sph), ph[z]);
        eg += Hi[z] * (-c * w.y);
        // (Ubg k + Vbg l) Re(i conj(ph) Sph)
        const cd w2 = cmul(cconj(ph[z]), sph);
        ag += Hi[z] * (P.Ubg[z] * kk + pr[L.Vbg() + z] * ll) * (-w2.y);
      }
      if (A.slot[QGK_D_ENSGENSPEC] >= 0) add(QGK_D_ENSGENSPEC, 0, eg / H * s);
      if (A.slot[QGK_D_APEGENSPEC] >= 0) add(QGK_D_APEGENSPEC, 0, ag / H * s);
    }
    if (A.slot[QGK_D_ENSFRICTIONSPEC] >= 0)
      add(QGK_D_ENSFRICTIONSPEC, 0,
          P.rek * Hi[bot] / H * wv2 * re_conj_mul(qh[bot], ph[bot]) * s);
    if (A.slot[QGK_D_APEFLUX] >= 0) {  // two layers
      const cd d = csub(ph[0], ph[1]);
      const cd j = A.J3[(size_t)m * plane + pt];
      add(QGK_D_APEFLUX, 0,
          pr[L.rd2inv()] * pr[L.del1()] * pr[L.del2()] * re_conj_mul(j, d) * s);
    }
    if (A.slot[QGK_D_KEFLUX] >= 0) {
      const cd j0 = A.Jpxi[base], j1 = A.Jpxi[base + plane];
      add(QGK_D_KEFLUX, 0,
          (pr[L.del1()] * re_conj_mul(j0, ph[0]) + pr[L.del2()] * re_conj_mul(j1, ph[1])) * s);
    }
    if (A.pts) {  // APEgen (qg_model.py:250-262): columns 1..nk-3 count twice
      const cd x = {pr[L.del1()] * ph[0].x + pr[L.del2()] * ph[1].x,
                    pr[L.del1()] * ph[0].y + pr[L.del2()] * ph[1].y};
      const cd w = cmul(x, cconj(csub(ph[0], ph[1])));
      const double mult = (k >= 1 && k <= A.nk - 3) ? 2.0 : 1.0;
      A.pts[i] = mult * (-kk * w.y);
    }
    if (A.slot[QGK_D_KESPEC_MODAL] >= 0 || A.slot[QGK_D_PESPEC_MODAL] >= 0) {
      for (int z = 0; z < nz; ++z) {
        cd pn = {0.0, 0.0};
        for (int j = 0; j < nz; ++j) {
          const double c = pr[L.pinv() + z * nz + j];
          pn.x += c * ph[j].x;
          pn.y += c * ph[j].y;
        }
        if (A.slot[QGK_D_KESPEC_MODAL] >= 0) add(QGK_D_KESPEC_MODAL, z, wv2 * abs2(pn) * s);
        if (z >= 1 && A.slot[QGK_D_PESPEC_MODAL] >= 0)
          add(QGK_D_PESPEC_MODAL, z - 1, pr[L.kdi2() + z] * abs2(pn) * s);
      }
    }
    if (A.slot[QGK_D_APESPEC] >= 0) {
      double a = 0.0;
      for (int z = 0; z + 1 < nz; ++z) a += pr[L.f2gpi() + z] * abs2(csub(ph[z], ph[z + 1]));
      add(QGK_D_APESPEC, 0, a / H * s);
    }
    if (A.slot[QGK_D_KEFLUX_DIV] >= 0) {
      double a = 0.0;
      for (int z = 0; z < nz; ++z) a += Hi[z] * re_conj_mul(ph[z], A.Jpxi[base + z * plane]);
      add(QGK_D_KEFLUX_DIV, 0, a / H * s);
    }
    if (A.slot[QGK_D_APEFLUX_DIV] >= 0) {
      double a = 0.0;
      for (int z = 0; z < nz; ++z) a += Hi[z] * re_conj_mul(ph[z], A.J3[base + z * plane]);
      add(QGK_D_APEFLUX_DIV, 0, a / H * s);
    }
  }
}

// Real-space products for a Jacobian: P1 = u * x, P2 = v * x per plane.
// mode 0: x, u, v have nzx = nz planes per member.  mode 1 (two layers):
// one plane x per member, u -> del1 u0 + del2 u1, v likewise (qg_model.py:216-219).
__global__ void k_diag_products(const double* __restrict__ x, const double* __restrict__ u,
                                const double* __restrict__ v, const double* __restrict__ prm,
                                double* __restrict__ P1, double* __restrict__ P2, int M, int nz,
                                int npix, int mode, int plen) {
  const int nzx = mode == 1 ? 1 : nz;
  const size_t total = (size_t)M * nzx * npix;
  const DiagLayout L{nz};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(i % npix);
    const int z = (int)((i / npix) % nzx);
    const int m = (int)(i / ((size_t)npix * nzx));
    double uu, vv;
    if (mode == 1) {
      const double d1 = prm[(size_t)m * plen + L.del1()], d2 = prm[(size_t)m * plen + L.del2()];
      const size_t o = (size_t)m * nz * npix + p;
      uu = d1 * u[o] + d2 * u[o + npix];
      vv = d1 * v[o] + d2 * v[o + npix];
    } else {
      uu = u[i];
      vv = v[i];
    }
    P1[i] = uu * x[i];
    P2[i] = vv * x[i];
  }
}

// J = sign * (ik W1 + il W2) over nzx planes per member
__global__ void k_diag_jacobian(const cd* __restrict__ W1, const cd* __restrict__ W2,
                                const MemberParams* __restrict__ mp, cd* __restrict__ J, int M,
                                int nzx, int nl, int ldk, double sign) {
  const size_t plane = (size_t)nl * ldk;
  const size_t total = (size_t)M * nzx * plane;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pt = i % plane;
    const int m = (int)(i / (plane * nzx));
    const int l = (int)(pt / ldk), k = (int)(pt % ldk);
    const double kk = mp[m].kk[k], ll = mp[m].ll[l];
    const cd a = W1[i], b = W2[i];
    J[i] = {sign * (-kk * a.y - ll * b.y), sign * (kk * a.x + ll * b.x)};
  }
}

// Spectral inputs of the inverse transforms: mode 0 xih = -wv2 ph (nz planes),
// mode 1 ph0 - ph1 (one plane), mode 2 S ph (nz planes).
__global__ void k_diag_prepare(const cd* __restrict__ ph, const MemberParams* __restrict__ mp,
                               const double* __restrict__ prm, cd* __restrict__ out, int M,
                               int nz, int nl, int ldk, int mode, int plen) {
  const size_t plane = (size_t)nl * ldk;
  const int nzx = mode == 1 ? 1 : nz;
  const size_t total = (size_t)M * nzx * plane;
  const DiagLayout L{nz};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pt = i % plane;
    const int z = (int)((i / plane) % nzx);
    const int m = (int)(i / (plane * nzx));
    const size_t base = (size_t)m * nz * plane + pt;
    cd r;
    if (mode == 0) {
      const double w = mp[m].k2l2[pt];
      const cd p = ph[base + z * plane];
      r = {-w * p.x, -w * p.y};
    } else if (mode == 1) {
      r = csub(ph[base], ph[base + plane]);
    } else {
      r = {0.0, 0.0};
      for (int j = 0; j < nz; ++j) {
        const double Sij = prm[(size_t)m * plen + L.S() + z * nz + j];
        const cd p = ph[base + j * plane];
        r.x += Sij * p.x;
        r.y += Sij * p.y;
      }
    }
    out[i] = r;
  }
}

// Deterministic block reduction of one plane per block: out[b] = sum(in plane b)
// (mode 0), or sum(u^2 + v^2) over a real plane (mode 1).
__global__ __launch_bounds__(256) void k_diag_plane_sum(const double* __restrict__ a,
                                                        const double* __restrict__ b,
                                                        double* __restrict__ out, size_t n,
                                                        int mode) {
  __shared__ double red[256];
  const size_t off = (size_t)blockIdx.x * n;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) {
    const double x = a[off + i];
    acc += mode == 1 ? x * x + b[off + i] * b[off + i] : x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// Scalar diagnostics from the plane sums: EKE[m][z] += 0.5 sum_z / npix;
// EKEdiss[m] += Hi[-1]/H rek sum_bottom / npix; APEgen[m] += U rd^-2 del1 del2
// sum / M^2.  sums: [M][nz] (u^2+v^2) then [M] (APEgen points).
__global__ void k_diag_scalars(const double* __restrict__ sums, const double* __restrict__ prm,
                               const MemberParams* __restrict__ mp, double* __restrict__ accs,
                               int M, int nz, int plen, double npix, double invM2, int s_eke,
                               int s_ekediss, int s_apegen) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const DiagLayout L{nz};
  const double* pr = prm + (size_t)m * plen;
  if (s_eke >= 0)
    for (int z = 0; z < nz; ++z)
      accs[(size_t)(s_eke + z) * M + m] += 0.5 * (sums[(size_t)m * nz + z] / npix);
  if (s_ekediss >= 0)
    accs[(size_t)s_ekediss * M + m] +=
        pr[L.Hi() + nz - 1] / pr[L.H()] * mp[m].rek * (sums[(size_t)m * nz + nz - 1] / npix);
  if (s_apegen >= 0)
    accs[(size_t)s_apegen * M + m] += pr[L.Ushear()] * pr[L.rd2inv()] * pr[L.del1()] *
                                      pr[L.del2()] * sums[(size_t)M * nz + m] * invM2;
}

// out[j][pt] = sum over members of acc[slot + j][m][pt]
__global__ void k_diag_member_sum(const double* __restrict__ acc, double* __restrict__ out,
                                  int M, int nsl, size_t plane) {
  const size_t total = (size_t)nsl * plane;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t j = i / plane, pt = i % plane;
    double s = 0.0;
    for (int m = 0; m < M; ++m) s += acc[(j * M + m) * plane + pt];
    out[i] = s;
  }
}

// ==========================================================================
// Built-in closures (pyqg/parameterizations.py:132-289): spectral derivatives,
// transforms and pointwise products over all members at once
// ==========================================================================

// out = cx * (ik X) + cy * (il Y) over nzx planes per member (X or Y may be null)
__global__ void k_spec_ikil(const cd* __restrict__ X, const cd* __restrict__ Y,
                            cd* __restrict__ out, const MemberParams* __restrict__ mp, int M,
                            int nzx, int nl, int ldk, double cx, double cy) {
  const size_t plane = (size_t)nl * ldk;
  const size_t total = (size_t)M * nzx * plane;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pt = i % plane;
    const int m = (int)(i / (plane * nzx));
    const int l = (int)(pt / ldk), k = (int)(pt % ldk);
    cd r = {0.0, 0.0};
    if (X) {  // ik X
      const double kk = mp[m].kk[k];
      const cd x = X[i];
      r = {-kk * x.y, kk * x.x};
      r = {cx * r.x, cx * r.y};
    }
    if (Y) {
      const double ll = mp[m].ll[l];
      const cd y = Y[i];
      r = {r.x + cy * (-ll * y.y), r.y + cy * (ll * y.x)};
    }
    out[i] = r;
  }
}

// out = a X + b Y (complex planes)
__global__ void k_caxpby(const cd* __restrict__ X, const cd* __restrict__ Y, cd* __restrict__ out,
                         double a, double b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    out[i] = {a * X[i].x + b * Y[i].x, a * X[i].y + b * Y[i].y};
}

// out = sign * wv2^power * X (wv2 = k^2 + l^2, so lap = -wv2)
__global__ void k_spec_wv2(const cd* __restrict__ X, cd* __restrict__ out,
                           const MemberParams* __restrict__ mp, int M, int nz, int nl, int ldk,
                           int power, double sign) {
  const size_t plane = (size_t)nl * ldk;
  const size_t total = (size_t)M * nz * plane;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pt = i % plane;
    const int m = (int)(i / (plane * nz));
    double w = mp[m].k2l2[pt];
    double f = sign;
    for (int p = 0; p < power; ++p) f *= w;
    const cd x = X[i];
    out[i] = {f * x.x, f * x.y};
  }
}

// Smagorinsky eddy viscosity (parameterizations.py:174-181): from Sxx, Syy and
// 2 Sxy (R0, R1, R2) computes nu into R3 and nu*Sxx, nu*Syy, nu*Sxy in place.
__global__ void k_smag_nu(double* __restrict__ R0, double* __restrict__ R1,
                          double* __restrict__ R2, double* __restrict__ R3,
                          const double* __restrict__ dx, double cs, int M, int nz, int npix) {
  const size_t total = (size_t)M * nz * npix;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / ((size_t)nz * npix));
    const double sxx = R0[i], syy = R1[i], sxy = 0.5 * R2[i];
    const double c = cs * dx[m];
    const double nu = (c * c) * sqrt(2 * (sxx * sxx + syy * syy + 2 * (sxy * sxy)));
    R3[i] = nu;
    R0[i] = nu * sxx;
    R1[i] = nu * syy;
    R2[i] = nu * sxy;
  }
}

// ZannaBolton2020 products (parameterizations.py:270-280): from vx, vy, ux, uy
// (R0..R3): R0 = vort*stretch, R1 = vort*shear, R2 = vort^2 + shear^2 + stretch^2.
__global__ void k_zb_products(double* __restrict__ R0, double* __restrict__ R1,
                              double* __restrict__ R2, const double* __restrict__ R3,
                              size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double vx = R0[i], vy = R1[i], ux = R2[i], uy = R3[i];
    const double rv = vx - uy, sh = vx + uy, st = ux - vy;
    R0[i] = rv * st;
    R1[i] = rv * sh;
    R2[i] = rv * rv + sh * sh + st * st;
  }
}

// out = (a * dx^2) * b per member (backscatter's lap_lap_psi * dx^2 * nu)
__global__ void k_bs_product(const double* __restrict__ a, const double* __restrict__ b,
                             const double* __restrict__ dx, double* __restrict__ out, int M,
                             int nz, int npix) {
  const size_t total = (size_t)M * nz * npix;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / ((size_t)nz * npix));
    out[i] = (a[i] * (dx[m] * dx[m])) * b[i];
  }
}

// Deterministic per-plane sums of a*b (one block per plane)
__global__ __launch_bounds__(256) void k_plane_dot(const double* __restrict__ a,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ out, size_t n) {
  __shared__ double red[256];
  const size_t off = (size_t)blockIdx.x * n;
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < n; i += 256) acc += a[off + i] * b[off + i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// BackscatterBiharmonic forcing (parameterizations.py:226-236): dissipation =
// -R2, backscatter = -Cb llpsi * sum_z Hi <psi diss> / (sum_z Hi <psi llpsi> + eps)
// per member, from plane sums s_pd[m][z] = sum psi R2 and s_pl[m][z] = sum psi
// llpsi; out = dissipation + backscatter.
__global__ void k_bs_final(const double* __restrict__ R2, const double* __restrict__ llpsi,
                           const double* __restrict__ s_pd, const double* __restrict__ s_pl,
                           const double* __restrict__ Hi, double cb, double eps,
                           double* __restrict__ out, int M, int nz, int npix) {
  const size_t total = (size_t)M * nz * npix;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i / ((size_t)nz * npix));
    double num = 0.0, den = 0.0;
    for (int z = 0; z < nz; ++z) {
      num += Hi[(size_t)m * nz + z] * (-s_pd[(size_t)m * nz + z] / npix);
      den += Hi[(size_t)m * nz + z] * (s_pl[(size_t)m * nz + z] / npix);
    }
    const double diss = -R2[i];
    out[i] = diss + (-cb * llpsi[i]) * (num / (den + eps));
  }
}

// out = (accumulate ? out : 0) + weight * (scale * a)
__global__ void k_param_out(const double* __restrict__ a, double* __restrict__ out, double scale,
                            double weight, int accumulate, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double v = (scale * a[i]) * weight;
    out[i] = accumulate ? out[i] + v : v;
  }
}

// Plain streaming copy, one 16-byte element per lane over a grid that covers
// the buffer: the calibration kernel for HBM traffic counters and the measured
// copy bandwidth.  Workgroups are dispatched in address order, so the lines in
// flight at any moment form one moving window; this reaches a higher rate on
// HBM3E than a grid-stride loop, whose lanes touch pages across the whole
// buffer at once.
__global__ void k_stream_copy(const double2* __restrict__ in, double2* __restrict__ out,
                              size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i];
}

// Memory-only stand-ins for the two fused kernels: the same fields, the same
// access patterns, cache hints and bytes, no transforms.  Their times bound what
// the fused kernels can reach on this memory system.
//
// Row phase: three fields read and two written, each a sequential stream over
// the first ncols values of every row of stride ldk.
__global__ void k_probe_rows(const cd* __restrict__ a, const cd* __restrict__ b,
                             const cd* __restrict__ c, cd* __restrict__ d, cd* __restrict__ e,
                             size_t nrows, int ncols, int ldk) {
  const size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t row = g / (size_t)ncols;
  if (row < nrows) {
    const size_t i = row * ldk + (g - row * ncols);
    const cd x = ld_stream(a + i), y = ld_stream(b + i), z = ld_stream(c + i);
    st_stream(d + i, {x.x + y.x, x.y + y.y});
    st_stream(e + i, {z.x + x.x, z.y + x.y});
  }
}

// Column phase: a block per (8-column tile, member), lanes (column c = tid % 8,
// row t = tid / 8) over rows t + 32 s, every layer: UQm, VQm, qh_old, dqhdt_p
// and dqhdt_pp read, qh_new, dqhdt (over dqhdt_pp), Qm, Um and Vm written.
__global__ __launch_bounds__(256) void k_probe_cols(
    const cd* __restrict__ uq, const cd* __restrict__ vq, const cd* __restrict__ qo,
    const cd* __restrict__ p, cd* __restrict__ pp, cd* __restrict__ qn, cd* __restrict__ qm,
    cd* __restrict__ um, cd* __restrict__ vm, int ny, int ldk, int nz) {
  const int c = threadIdx.x % 8, t = threadIdx.x / 8;
  const int k = blockIdx.x * 8 + c;
  const size_t plane = (size_t)ny * ldk;
  for (int z = 0; z < nz; ++z) {
    const size_t base = ((size_t)blockIdx.y * nz + z) * plane + k;
    for (int r0 = 0; r0 < ny; r0 += 256) {
      cd v[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int row = r0 + t + 32 * s;
        v[s] = cd{0.0, 0.0};
        if (row < ny) {
          const size_t o = base + (size_t)row * ldk;
          const cd a = ld_stream(uq + o), b = ld_stream(vq + o), q = ld_stream(qo + o),
                   d = ld_stream(p + o), e = ld_stream(pp + o);
          v[s] = {a.x + b.x + q.x + d.x + e.x, a.y + b.y + q.y + d.y + e.y};
        }
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int row = r0 + t + 32 * s;
        if (row < ny) {
          const size_t o = base + (size_t)row * ldk;
          st_stream(qn + o, v[s]);
          st_stream(pp + o, v[s]);
          st_stream(qm + o, v[s]);
          st_stream(um + o, v[s]);
          st_stream(vm + o, v[s]);
        }
      }
    }
  }
}

// ==========================================================================
// host side
// ==========================================================================

std::vector<cd> make_twiddles(int n) {
  std::vector<cd> tw(n);
  const long double two_pi = 6.283185307179586476925286766559005768L;
  for (int m = 0; m < n; ++m) {
    long double th = two_pi * (long double)m / (long double)n;
    tw[m] = {(double)cosl(th), (double)(-sinl(th))};
  }
  return tw;
}

// Grid sizes: 2^a in [8, 1024] or 3 * 2^a in [24, 768].
bool size_supported(int n) {
  const int m = n % 3 == 0 ? n / 3 : n;
  const bool pow2 = m > 0 && (m & (m - 1)) == 0;
  return pow2 && (n % 3 == 0 ? (n >= 24 && n <= 768) : (n >= 8 && n <= 1024));
}

}  // namespace

struct qgk_bucket {
  int M = 0, nz = 0, ny = 0, nx = 0, nl = 0, nk = 0, flags = 0;
  int ldk = 0;  // row stride of spectral arrays (nk padded to a multiple of 8)
  size_t nplane_c = 0, nplane_r = 0;  // points per spectral / real plane
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  // state
  cd* field_c[QGK_NFIELDS] = {};
  double* field_r[QGK_NFIELDS] = {};
  cd* hist[3] = {};  // rotating: hist[iD] dqhdt, hist[iP] dqhdt_p, hist[iPP] dqhdt_pp
  cd* qh_alt = nullptr;  // the fused step writes qh here, then the two swap
  int iD = 0, iP = 1, iPP = 2;
  cd *Qm = nullptr, *Um = nullptr, *Vm = nullptr, *UQm = nullptr, *VQm = nullptr;
  cd* phm = nullptr;  // ph of the current state for the fused step, 3 <= nz <= 16
  cd* ctmp = nullptr;      // spectral scratch (M*nz planes)
  double* rtmp = nullptr;  // real scratch (M*nz planes)
  bool primed = false;     // Qm/Um/Vm match the current qh
  bool q_stale = false;    // the real q lags qh (after fused steps)
  // constants
  std::vector<MemberParams> hp;
  MemberParams* dp = nullptr;
  bool params_dirty = true;
  double* shared_tab[QGK_NCONSTS] = {};
  std::vector<std::vector<double*>> own_tab;  // [which][member]
  cd* tw_x = nullptr;
  cd* tw_y = nullptr;
  double* tri_tab = nullptr;  // [M][3][nz] bands of S for the tridiagonal solve
  // timing
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;
  struct Pending {
    hipEvent_t a, b;
    int kind;
  };
  std::vector<Pending> pending;
  double ms_row = 0, ms_col = 0;
  long n_row = 0, n_col = 0;
  // graph of one fused step (per parity of the history rotation)
  hipGraphExec_t gexec[3] = {};
  StepCoeffs gcoef = {0, 0, 0};
  bool gvalid = false;
  // register radix of the fused row and column kernels (autotuned)
  int erow = 8, ecol = 8;
  bool csplit = false;  // nz = 2: column steps by k_col_step_split (one member)
  bool pk_off = false;  // packed x-mixed exchange switched off (qgk_set_packed_exchange)
  bool grid_std = false;  // every member's ll and k2l2 tables are standard (k_grid_check)
  int* gflag = nullptr;   // [M] per-member result of k_grid_check
  unsigned long long* speed = nullptr;  // [M] member health words (qgk_member_speed)
  double* small = nullptr;  // 256 doubles: learned-closure scales (input [0,128), output [128,256))
  std::vector<double> small_in_h, small_out_h;  // values last uploaded there
  // diagnostics: per-member parameters, accumulators and work space
  std::vector<double> dprm_h;  // [M][plen]
  std::vector<char> dprm_set;  // per member: parameters given
  double* dprm = nullptr;
  bool dprm_dirty = true;
  int dslot[QGK_NDIAG] = {};   // first slot (spectral or scalar set), -1 = invalid for nz
  int dnsl[QGK_NDIAG] = {};    // slots per diagnostic
  int nslot = 0, nscal = 0;
  double *dacc = nullptr, *dsc = nullptr;      // running sums [slot][M][nl][ldk], [slot][M]
  double *dacc_i = nullptr, *dsc_i = nullptr;  // one-shot evaluation
  cd* dW = nullptr;            // 5 spectral sets of M*nz planes: W1, W2, Jq, Jpxi, J3
  double* dR = nullptr;        // 3 real sets of M*nz planes: x, P1, P2
  double* dsum = nullptr;      // M*nz + M plane sums; then nz planes of member sums
  // built-in closures: per-member grid spacing and layer thicknesses, plane sums
  double *pdx = nullptr, *pHi = nullptr, *psum = nullptr;
  std::vector<double> pdx_h, pHi_h;  // last values uploaded to pdx, pHi
  // user step graph (qgk_capture_begin / qgk_capture_end / qgk_graph_replay)
  bool capturing = false;
  hipGraphExec_t ugexec = nullptr;
  int ug_phase = -1;  // history rotation (iP) at capture
  long ug_gen = -1;   // gen at capture
  const cd* ug_qh = nullptr;  // qh buffer at capture (the fused step swaps qh buffers)
  bool ug_fused = false;      // the captured steps are fused closed steps (qgk_fused_*)
  // stochastic forcing (qgk_set_spectral_noise): AR(1) state, amplitudes, step counter
  cd* noise_xi = nullptr;
  double* noise_amp = nullptr;
  unsigned long long* noise_ctr = nullptr;
  bool noise_on = false;
  double noise_phi = 0.0;
  unsigned long long noise_seed = 0;
  hipStream_t lst = nullptr;  // stream override for fused-step launches (graph capture)
  long gen = 0;               // bumped whenever launch arguments change (graph cache key)
  // hipGraph of two fused steps of a set of buckets, replayed while the AB
  // weights are constant (owned by the first bucket of the set)
  struct GraphCache {
    std::vector<qgk_bucket*> bs;
    std::vector<long> gens;
    std::vector<cd*> qh0;
    std::vector<int> iP0;
    StepCoeffs cf = {0, 0, 0};
    hipGraphExec_t exec = nullptr;
  } graph;
};

namespace {

int dev_alloc(qgk_bucket* b, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess)
    return fail(QGK_E_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  b->allocs.push_back(*p);
  e = hipMemsetAsync(*p, 0, bytes, b->stream);
  if (e != hipSuccess) return fail(QGK_E_HIP, "hipMemsetAsync failed: %s", hipGetErrorString(e));
  return 0;
}

bool is_real_field(int f) {
  return f == QGK_Q || f == QGK_U || f == QGK_V || f == QGK_UQ || f == QGK_VQ || f == QGK_DQ ||
         f == QGK_DU || f == QGK_DV;
}

size_t const_len(const qgk_bucket* b, int which) {
  switch (which) {
    case QGK_KK: return b->ldk;
    case QGK_LL: return b->nl;
    case QGK_K2L2:
    case QGK_FILTR: return b->nplane_c;
    case QGK_A: return (size_t)b->nz * b->nz * b->nplane_c;
    case QGK_UBG:
    case QGK_QY: return b->nz;
  }
  return 0;
}

const double*& param_slot(MemberParams& p, int which) {
  switch (which) {
    case QGK_KK: return p.kk;
    case QGK_LL: return p.ll;
    case QGK_K2L2: return p.k2l2;
    case QGK_FILTR: return p.filtr;
    case QGK_A: return p.a;
    case QGK_UBG: return p.Ubg;
    default: return p.Qy;
  }
}

// Entry points that read or write host memory, synchronize, or change a
// constant refuse to run while the bucket's stream is being captured: a copy
// from pageable host memory recorded into a graph would read that host buffer
// again at every replay, after it may be gone, and a synchronization is
// illegal inside a capture.
int refuse_in_capture(const qgk_bucket* b, const char* what) {
  if (b->capturing) return fail(QGK_E_STATE, "%s is not allowed during graph capture", what);
  return 0;
}

int upload_params(qgk_bucket* b) {
  if (!b->params_dirty) return 0;
  if (b->capturing)
    return fail(QGK_E_STATE, "member constants changed during graph capture");
  HIPC(hipMemcpyAsync(b->dp, b->hp.data(), sizeof(MemberParams) * b->M, hipMemcpyHostToDevice,
                      b->stream));
  b->params_dirty = false;
  if (b->nz <= 2) {  // the packed register kernels form ll and kappa^2 on a standard grid
    if (!b->gflag) {
      int rc = dev_alloc(b, (void**)&b->gflag, b->M * sizeof(int));
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_grid_check, dim3(b->M), dim3(256), 0, b->stream, b->dp, b->ny, b->nk,
                       b->ldk, b->gflag);
    HIPC(hipGetLastError());
    std::vector<int> f(b->M);
    HIPC(hipMemcpyAsync(f.data(), b->gflag, b->M * sizeof(int), hipMemcpyDeviceToHost,
                        b->stream));
    HIPC(hipStreamSynchronize(b->stream));
    bool all = true;
    for (int v : f) all = all && v == 1;
    if (all != b->grid_std) {  // the exchange format may change with it
      b->grid_std = all;
      b->primed = false;
      b->gvalid = false;
      b->gen++;
    }
  }
  return 0;
}

// Device address of `scratch` on the current device, looked up once per
// device (each device holds its own copy of the array).
constexpr int kMaxDevices = 64;
unsigned long long* g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_address(unsigned long long** out) {
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(QGK_E_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(QGK_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_dev[dev] = static_cast<unsigned long long*>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

cd* D(qgk_bucket* b) { return b->hist[b->iD]; }
cd* Pp(qgk_bucket* b) { return b->hist[b->iP]; }
cd* PP(qgk_bucket* b) { return b->hist[b->iPP]; }

// ---- event timing
int t_begin(qgk_bucket* b, hipEvent_t* ev) {
  if (!b->timing) return 0;
  HIPC(hipEventCreate(ev));
  HIPC(hipEventRecord(*ev, b->stream));
  return 0;
}
int t_end(qgk_bucket* b, hipEvent_t ev0, int kind) {
  if (!b->timing) return 0;
  hipEvent_t ev1;
  HIPC(hipEventCreate(&ev1));
  HIPC(hipEventRecord(ev1, b->stream));
  b->pending.push_back({ev0, ev1, kind});
  return 0;
}
int t_collect(qgk_bucket* b) {
  for (auto& p : b->pending) {
    HIPC(hipEventSynchronize(p.b));
    float ms = 0;
    HIPC(hipEventElapsedTime(&ms, p.a, p.b));
    if (p.kind == 0) {
      b->ms_row += ms;
      b->n_row++;
    } else {
      b->ms_col += ms;
      b->n_col++;
    }
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  b->pending.clear();
  return 0;
}

// ---- dispatch helpers
int grid_1d(size_t total) {
  size_t g = (total + 255) / 256;
  if (g > 65535 * 8) g = 65535 * 8;
  return (int)(g ? g : 1);
}

// Launches over nf fields of nplanes planes each (block rows f * nplanes + p).
template <int SIGN>
int launch_col_fft_n(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes) {
  const int ny = b->ny, nk = b->nk;
#define QGK_COL(NY)                                                                     \
  case NY: {                                                                            \
    using G = Geo<NY>;                                                                  \
    dim3 grid((nk + G::TK - 1) / G::TK, nf * nplanes);                                  \
    hipLaunchKernelGGL((k_col_fft<NY, SIGN>), grid, dim3(G::T * G::TK), 0, b->stream, j,  \
                       nplanes, nk, b->ldk, b->tw_y);                                   \
    break;                                                                              \
  }
  switch (ny) {
    QGK_COL(8) QGK_COL(16) QGK_COL(32) QGK_COL(64) QGK_COL(128) QGK_COL(256) QGK_COL(512)
    QGK_COL(1024) QGK_COL(24) QGK_COL(48) QGK_COL(96) QGK_COL(192) QGK_COL(384) QGK_COL(768)
    default: return fail(QGK_E_ARG, "unsupported ny=%d", ny);
  }
#undef QGK_COL
  HIPC(hipGetLastError());
  return 0;
}

// pk: the input rows are the fused register path's packed x-mixed rows
int launch_row_c2r_n(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes, bool pk = false) {
  const int nx = b->nx, ny = b->ny;
  const double scale = 1.0 / ((double)nx * (double)ny);
#define QGK_ROW(NX)                                                                       \
  case NX: {                                                                              \
    using G = Geo<NX>;                                                                    \
    dim3 grid((ny / 2 + G::RP - 1) / G::RP, nf * nplanes);                                \
    if (pk)                                                                               \
      hipLaunchKernelGGL((k_row_c2r<NX, true>), grid, dim3(G::T * G::RP), 0, b->stream, j, \
                         nplanes, ny, scale, b->tw_x);                                    \
    else                                                                                  \
      hipLaunchKernelGGL((k_row_c2r<NX>), grid, dim3(G::T * G::RP), 0, b->stream, j,      \
                         nplanes, ny, scale, b->tw_x);                                    \
    break;                                                                                \
  }
  switch (nx) {
    QGK_ROW(8) QGK_ROW(16) QGK_ROW(32) QGK_ROW(64) QGK_ROW(128) QGK_ROW(256) QGK_ROW(512)
    QGK_ROW(1024) QGK_ROW(24) QGK_ROW(48) QGK_ROW(96) QGK_ROW(192) QGK_ROW(384) QGK_ROW(768)
    default: return fail(QGK_E_ARG, "unsupported nx=%d", nx);
  }
#undef QGK_ROW
  HIPC(hipGetLastError());
  return 0;
}

int launch_row_r2c_n(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes) {
  const int nx = b->nx, ny = b->ny;
#define QGK_ROW(NX)                                                                        \
  case NX: {                                                                               \
    using G = Geo<NX>;                                                                     \
    dim3 grid((ny / 2 + G::RP - 1) / G::RP, nf * nplanes);                                 \
    hipLaunchKernelGGL((k_row_r2c<NX>), grid, dim3(G::T * G::RP), 0, b->stream, j, nplanes,  \
                       ny, b->tw_x);                                                       \
    break;                                                                                 \
  }
  switch (nx) {
    QGK_ROW(8) QGK_ROW(16) QGK_ROW(32) QGK_ROW(64) QGK_ROW(128) QGK_ROW(256) QGK_ROW(512)
    QGK_ROW(1024) QGK_ROW(24) QGK_ROW(48) QGK_ROW(96) QGK_ROW(192) QGK_ROW(384) QGK_ROW(768)
    default: return fail(QGK_E_ARG, "unsupported nx=%d", nx);
  }
#undef QGK_ROW
  HIPC(hipGetLastError());
  return 0;
}

PlaneJobs one_job(const void* in, void* out) {
  PlaneJobs j{};
  j.in[0] = in;
  j.out[0] = out;
  return j;
}

template <int SIGN>
int launch_col_fft(qgk_bucket* b, const cd* in, cd* out, int nplanes) {
  return launch_col_fft_n<SIGN>(b, one_job(in, out), 1, nplanes);
}

int launch_row_c2r(qgk_bucket* b, const cd* in, double* out, int nplanes, bool pk = false) {
  return launch_row_c2r_n(b, one_job(in, out), 1, nplanes, pk);
}

int launch_row_r2c(qgk_bucket* b, const double* in, cd* out, int nplanes) {
  return launch_row_r2c_n(b, one_job(in, out), 1, nplanes);
}

// one-launch plane transforms: square grids of n <= kPlaneMax
bool plane_ok(const qgk_bucket* b) { return b->nx == b->ny && b->nx <= kPlaneMax; }

struct PlaneR2CL {
  template <int N>
  static void go(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes) {
    hipLaunchKernelGGL((k_plane_r2c<N>), dim3(nf * nplanes), dim3(PlaneGeo<N>::THREADS), 0,
                       b->stream, j, nplanes, b->ldk, b->tw_x, b->tw_y);
  }
};

struct PlaneC2RL {
  template <int N>
  static void go(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes) {
    const double scale = 1.0 / ((double)N * (double)N);
    hipLaunchKernelGGL((k_plane_c2r<N, false>), dim3(nf * nplanes), dim3(PlaneGeo<N>::THREADS),
                       0, b->stream, j, SpecJobs{}, b->dp, b->nz, nplanes, b->ldk, scale,
                       b->tw_x, b->tw_y);
  }
};

struct PlaneSpecL {
  template <int N>
  static void go(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes, const SpecJobs& dj) {
    const double scale = 1.0 / ((double)N * (double)N);
    hipLaunchKernelGGL((k_plane_c2r<N, true>), dim3(nf * nplanes), dim3(PlaneGeo<N>::THREADS),
                       0, b->stream, j, dj, b->dp, b->nz, nplanes, b->ldk, scale, b->tw_x,
                       b->tw_y);
  }
};

template <class L, class... A>
int launch_plane(qgk_bucket* b, const PlaneJobs& j, int nf, int nplanes, A... a) {
  switch (b->nx) {
    case 8: L::template go<8>(b, j, nf, nplanes, a...); break;
    case 16: L::template go<16>(b, j, nf, nplanes, a...); break;
    case 24: L::template go<24>(b, j, nf, nplanes, a...); break;
    case 32: L::template go<32>(b, j, nf, nplanes, a...); break;
    case 48: L::template go<48>(b, j, nf, nplanes, a...); break;
    case 64: L::template go<64>(b, j, nf, nplanes, a...); break;
    default: return fail(QGK_E_ARG, "no one-launch transform for n=%d", b->nx);
  }
  HIPC(hipGetLastError());
  return 0;
}

// true when some field's output bytes overlap some field's input bytes
bool jobs_overlap(int nf, const void* const* in, size_t in_bytes, void* const* out,
                  size_t out_bytes) {
  for (int i = 0; i < nf; ++i)
    for (int k = 0; k < nf; ++k) {
      const char* a = static_cast<const char*>(out[i]);
      const char* c = static_cast<const char*>(in[k]);
      if (a < c + in_bytes && c < a + out_bytes) return true;
    }
  return false;
}

// 2-D transforms of nf fields of nplanes planes each: rfftn / irfftn over the
// last two axes.  Small square grids take one launch for all fields; the rest
// a row launch and a column launch per field.
int r2c2d_n(qgk_bucket* b, int nf, const double* const* in, cd* const* out, int nplanes) {
  if (nf < 1 || nf > kPlaneJobs) return fail(QGK_E_ARG, "1 to %d fields per transform", kPlaneJobs);
  const size_t ib = (size_t)nplanes * b->nplane_r * sizeof(double);
  const size_t ob = (size_t)nplanes * b->nplane_c * sizeof(cd);
  if (plane_ok(b) &&
      !jobs_overlap(nf, reinterpret_cast<const void* const*>(in), ib,
                    reinterpret_cast<void* const*>(out), ob)) {
    PlaneJobs j{};
    for (int i = 0; i < nf; ++i) {
      j.in[i] = in[i];
      j.out[i] = out[i];
    }
    return launch_plane<PlaneR2CL>(b, j, nf, nplanes);
  }
  if ((long)nf * nplanes <= 65535 &&
      !jobs_overlap(nf, reinterpret_cast<const void* const*>(in), ib,
                    reinterpret_cast<void* const*>(out), ob)) {
    // one row launch and one column launch (in place on the outputs) for all fields
    PlaneJobs j{}, jc{};
    for (int i = 0; i < nf; ++i) {
      j.in[i] = in[i];
      j.out[i] = out[i];
      jc.in[i] = out[i];
      jc.out[i] = out[i];
    }
    int rc = launch_row_r2c_n(b, j, nf, nplanes);
    if (rc) return rc;
    return launch_col_fft_n<-1>(b, jc, nf, nplanes);
  }
  for (int i = 0; i < nf; ++i) {
    int rc = launch_row_r2c(b, in[i], out[i], nplanes);
    if (rc) return rc;
    if ((rc = launch_col_fft<-1>(b, out[i], out[i], nplanes))) return rc;
  }
  return 0;
}

int c2r2d_n(qgk_bucket* b, int nf, const cd* const* in, double* const* out, int nplanes) {
  if (nf < 1 || nf > kPlaneJobs) return fail(QGK_E_ARG, "1 to %d fields per transform", kPlaneJobs);
  const size_t ib = (size_t)nplanes * b->nplane_c * sizeof(cd);
  const size_t ob = (size_t)nplanes * b->nplane_r * sizeof(double);
  if (plane_ok(b) &&
      !jobs_overlap(nf, reinterpret_cast<const void* const*>(in), ib,
                    reinterpret_cast<void* const*>(out), ob)) {
    PlaneJobs j{};
    for (int i = 0; i < nf; ++i) {
      j.in[i] = in[i];
      j.out[i] = out[i];
    }
    return launch_plane<PlaneC2RL>(b, j, nf, nplanes);
  }
  // y pass into spectral scratch, x pass into out
  if (nplanes > b->M * b->nz) return fail(QGK_E_ARG, "too many planes for scratch");
  if (nf <= 3 && (long)nf * nplanes <= 65535) {
    // one column launch and one row launch for all fields; the scratch of
    // fields 1 and 2 is the fused step's x-mixed flux pair, which the fused
    // row kernel rewrites before anything reads it
    cd* stage[3] = {b->ctmp, b->UQm, b->VQm};
    PlaneJobs jc{}, jr{};
    for (int i = 0; i < nf; ++i) {
      jc.in[i] = in[i];
      jc.out[i] = stage[i];
      jr.in[i] = stage[i];
      jr.out[i] = out[i];
    }
    int rc = launch_col_fft_n<+1>(b, jc, nf, nplanes);
    if (rc) return rc;
    return launch_row_c2r_n(b, jr, nf, nplanes);
  }
  for (int i = 0; i < nf; ++i) {
    int rc = launch_col_fft<+1>(b, in[i], b->ctmp, nplanes);
    if (rc) return rc;
    if ((rc = launch_row_c2r(b, b->ctmp, out[i], nplanes))) return rc;
  }
  return 0;
}

int r2c2d(qgk_bucket* b, const double* in, cd* out, int nplanes) {
  return r2c2d_n(b, 1, &in, &out, nplanes);
}

int c2r2d(qgk_bucket* b, const cd* in, double* out, int nplanes) {
  return c2r2d_n(b, 1, &in, &out, nplanes);
}

// Calls L::go<N, E>(args...) for a transform length n (2^a or 3 * 2^a) and
// register radix choice e (4, 8 or 16: clamped to n for powers of two, 6 or 12
// for lengths with the factor 3).
template <class L, class... A>
int with_ne(int n, int e, A... a) {
#define QGK_NE(N)                                                               \
  case N:                                                                       \
    switch (e) {                                                                \
      case 4: L::template go<N, RadixFor<N, 4>::value>(a...); return 0;         \
      case 8: L::template go<N, RadixFor<N, 8>::value>(a...); return 0;         \
      case 16: L::template go<N, RadixFor<N, 16>::value>(a...); return 0;       \
      default: return fail(QGK_E_ARG, "unsupported radix %d", e);               \
    }
  switch (n) {
    QGK_NE(8) QGK_NE(16) QGK_NE(32) QGK_NE(64) QGK_NE(128) QGK_NE(256) QGK_NE(512)
    QGK_NE(1024) QGK_NE(24) QGK_NE(48) QGK_NE(96) QGK_NE(192) QGK_NE(384) QGK_NE(768)
    default: return fail(QGK_E_ARG, "unsupported transform length %d", n);
  }
#undef QGK_NE
}

// The packed register column kernel: six LDS columns of the transform buffer
// for its exchange, and its LDS (transform buffer, twiddles and the Nyquist
// column's q, u, v spectra of every layer) within the CU's 160 KB
template <int NY, int NZ, int E>
constexpr bool pk_fits() {
  using G = FGeo<NY, E>;
  return G::TK >= 6 && (NY * G::TK + NY + 3 * NZ * NY) * (int)sizeof(cd) <= 160 * 1024;
}

struct PackedL {
  template <int NY, int E>
  static void go(const qgk_bucket* b, bool* pk) {
    const bool split = b->nz == 2 && b->csplit && split_fits<NY, E>();
    *pk = (b->nz == 1 && pk_fits<NY, 1, E>()) || (b->nz == 2 && !split && pk_fits<NY, 2, E>());
  }
};

// The register column kernels (nz <= 2, unless the layer-split kernel steps
// the bucket) and the row kernel exchange packed x-mixed rows: NX/2 slots of
// whole lines per row, bins 0 and nx/2 sharing slot 0 (k_row_flux).
bool packed_exchange(const qgk_bucket* b) {
  if (b->nz > 2 || b->pk_off || !b->grid_std) return false;
  bool pk = false;
  if (with_ne<PackedL>(b->ny, b->ecol, b, &pk)) return false;
  return pk;
}

struct RowFluxL {
  template <int NX, int E>
  static void go(qgk_bucket* b) {
    using G = FGeo<NX, E>;
    const double scale = 1.0 / ((double)b->nx * (double)b->ny);
    dim3 grid((b->ny / 2 + G::RP - 1) / G::RP, b->M * b->nz);
    if (packed_exchange(b))
      hipLaunchKernelGGL((k_row_flux<NX, E, true>), grid, dim3(G::T * G::RP), 0,
                         b->lst ? b->lst : b->stream, b->Qm, b->Um, b->Vm, b->UQm, b->VQm, b->dp,
                         b->nz, b->ny, scale, b->tw_x);
    else
      hipLaunchKernelGGL((k_row_flux<NX, E, false>), grid, dim3(G::T * G::RP), 0,
                         b->lst ? b->lst : b->stream, b->Qm, b->Um, b->Vm, b->UQm, b->VQm, b->dp,
                         b->nz, b->ny, scale, b->tw_x);
  }
};

// k_col_prologue_reg / k_col_step_reg of NZ layers, packed when pk (and the
// packed kernel fits)
template <int NY, int NZ, int E>
void launch_prologue_reg(qgk_bucket* b, bool pk) {
  using G = FGeo<NY, E>;
  const dim3 block(G::T * G::TK);
  const cd* qh = b->field_c[QGK_QH];
  if constexpr (pk_fits<NY, NZ, E>()) {
    if (pk) {
      dim3 grid((col_count(b->nk, true) + G::TK - 1) / G::TK, b->M);
      hipLaunchKernelGGL((k_col_prologue_reg<NY, NZ, E, true>), grid, block, 0, b->stream, qh,
                         b->Qm, b->Um, b->Vm, b->dp, b->nk, b->ldk, b->tw_y);
      return;
    }
  }
  dim3 grid((b->nk + G::TK - 1) / G::TK, b->M);
  hipLaunchKernelGGL((k_col_prologue_reg<NY, NZ, E, false>), grid, block, 0, b->stream, qh,
                     b->Qm, b->Um, b->Vm, b->dp, b->nk, b->ldk, b->tw_y);
}

template <int NY, int NZ, int E>
void launch_step_reg(qgk_bucket* b, bool pk, StepCoeffs cf) {
  using G = FGeo<NY, E>;
  const dim3 block(G::T * G::TK);
  hipStream_t st = b->lst ? b->lst : b->stream;
  const cd* qo = b->field_c[QGK_QH];
  cd* qn = b->qh_alt;
  if constexpr (pk_fits<NY, NZ, E>()) {
    if (pk) {
      dim3 grid((col_count(b->nk, true) + G::TK - 1) / G::TK, b->M);
      hipLaunchKernelGGL((k_col_step_reg<NY, NZ, E, true>), grid, block, 0, st, qo, qn, Pp(b),
                         PP(b), b->UQm, b->VQm, b->Qm, b->Um, b->Vm, b->dp, cf, b->nk, b->ldk,
                         b->tw_y);
      return;
    }
  }
  dim3 grid((b->nk + G::TK - 1) / G::TK, b->M);
  hipLaunchKernelGGL((k_col_step_reg<NY, NZ, E, false>), grid, block, 0, st, qo, qn, Pp(b), PP(b),
                     b->UQm, b->VQm, b->Qm, b->Um, b->Vm, b->dp, cf, b->nk, b->ldk, b->tw_y);
}

struct ColPrologueL {
  template <int NY, int E>
  static void go(qgk_bucket* b) {
    using G = FGeo<NY, E>;
    dim3 grid((b->nk + G::TK - 1) / G::TK, b->M), block(G::T * G::TK);
    const cd* qh = b->field_c[QGK_QH];
    const bool pk = packed_exchange(b);
    if (b->nz == 1)
      launch_prologue_reg<NY, 1, E>(b, pk);
    else if (b->nz == 2)
      launch_prologue_reg<NY, 2, E>(b, pk);
    else if (b->phm)
      hipLaunchKernelGGL((k_col_prologue_ph<NY, E>), grid, block, 0, b->stream, qh, b->phm,
                         b->Qm, b->Um, b->Vm, b->dp, b->nz, b->nk, b->ldk, b->tw_y);
    else
      hipLaunchKernelGGL((k_col_prologue<NY, E>), grid, block, 0, b->stream, qh, b->Qm, b->Um,
                         b->Vm, b->dp, b->nz, b->nk, b->ldk, b->tw_y);
  }
};

struct ColStepL {
  template <int NY, int E>
  static void go(qgk_bucket* b, StepCoeffs cf) {
    using G = FGeo<NY, E>;
    dim3 grid((b->nk + G::TK - 1) / G::TK, b->M), block(G::T * G::TK);
    const cd* qo = b->field_c[QGK_QH];
    cd* qn = b->qh_alt;
    const bool pk = packed_exchange(b);
    if (b->nz == 1)
      launch_step_reg<NY, 1, E>(b, pk, cf);
    else if (b->nz == 2 && b->csplit && split_fits<NY, E>()) {
      if constexpr (split_fits<NY, E>())
        hipLaunchKernelGGL((k_col_step_split<NY, E>), grid, dim3(2 * G::T * G::TK), 0,
                           b->lst ? b->lst : b->stream, qo, qn, Pp(b), PP(b), b->UQm, b->VQm,
                           b->Qm, b->Um, b->Vm, b->dp, cf, b->nk, b->ldk, b->tw_y);
    } else if (b->nz == 2)
      launch_step_reg<NY, 2, E>(b, pk, cf);
    else if (b->phm)
      hipLaunchKernelGGL((k_col_step_ph<NY, E>), grid, block, 0, b->lst ? b->lst : b->stream, qo,
                         qn, Pp(b), PP(b), b->UQm, b->VQm, b->phm, b->Qm, b->Um, b->Vm, b->dp, cf,
                         b->nz, b->nk, b->ldk, b->tw_y);
    else
      hipLaunchKernelGGL((k_col_step<NY, E>), grid, block, 0, b->lst ? b->lst : b->stream, qo, qn, Pp(b), PP(b),
                         b->UQm, b->VQm, b->Qm, b->Um, b->Vm, b->dp, cf, b->nz, b->nk, b->ldk,
                         b->tw_y);
  }
};

int launch_row_flux(qgk_bucket* b) {
  int rc = with_ne<RowFluxL>(b->nx, b->erow, b);
  if (rc) return rc;
  HIPC(hipGetLastError());
  return 0;
}

int launch_col_prologue(qgk_bucket* b) {
  int rc = with_ne<ColPrologueL>(b->ny, b->ecol, b);
  if (rc) return rc;
  HIPC(hipGetLastError());
  return 0;
}

int launch_col_step(qgk_bucket* b, StepCoeffs cf) {
  int rc = with_ne<ColStepL>(b->ny, b->ecol, b, cf);
  if (rc) return rc;
  HIPC(hipGetLastError());
  return 0;
}

// Fused closed step (qgk_fused_row_net / qgk_fused_col_forcing): the packed
// register kernels with the network input written by the row kernel and the
// forcing added by the column kernel, for two-layer grids up to 128.
constexpr int kFusedNetMaxN = 128;

struct RowFluxNetL {
  template <int NX, int E>
  static void go(qgk_bucket* b, const double* sc, unsigned short* net_in, int cinp) {
    if constexpr (NX <= kFusedNetMaxN) {
      using G = FGeo<NX, E>;
      const double scale = 1.0 / ((double)b->nx * (double)b->ny);
      dim3 grid((b->ny / 2 + G::RP - 1) / G::RP, b->M * b->nz);
      hipLaunchKernelGGL((k_row_flux<NX, E, true, true>), grid, dim3(G::T * G::RP), 0,
                         b->stream, b->Qm, b->Um, b->Vm, b->UQm, b->VQm, b->dp, b->nz, b->ny,
                         scale, b->tw_x, sc, net_in, cinp);
    }
  }
};

struct ColStepForceL {
  template <int NY, int E>
  static void go(qgk_bucket* b, StepCoeffs cf) {
    if constexpr (NY <= kFusedNetMaxN && pk_fits<NY, 2, E>()) {
      using G = FGeo<NY, E>;
      dim3 grid((col_count(b->nk, true) + G::TK - 1) / G::TK, b->M);
      hipLaunchKernelGGL((k_col_step_reg<NY, 2, E, true, true>), grid, dim3(G::T * G::TK), 0,
                         b->stream, b->field_c[QGK_QH], b->qh_alt, Pp(b), PP(b), b->UQm, b->VQm,
                         b->Qm, b->Um, b->Vm, b->dp, cf, b->nk, b->ldk, b->tw_y,
                         b->field_c[QGK_DQH]);
    }
  }
};

// Members' health words into w[M] (cleared first, by a fill kernel, so the
// work can be captured): x inverse transforms of the primed Qm, Um, Vm into
// the real scratch, each folded into the words.
int launch_member_speed(qgk_bucket* b, unsigned long long* w) {
  int rc = upload_params(b);
  if (rc) return rc;
  if (!b->primed) {
    if ((rc = launch_col_prologue(b))) return rc;
    b->primed = true;
  }
  hipLaunchKernelGGL(k_fill_u64, dim3((b->M + 255) / 256), dim3(256), 0, b->stream, w, b->M,
                     0ULL);
  const int np = b->M * b->nz;
  const int parts = (int)std::min<size_t>(8, (b->nplane_r + 2047) / 2048);
  const cd* src[3] = {b->Um, b->Vm, b->Qm};
  const bool pk = packed_exchange(b);
  for (int f = 0; f < 3; ++f) {
    if ((rc = launch_row_c2r(b, src[f], b->rtmp, np, pk))) return rc;
    hipLaunchKernelGGL(k_speed_fold, dim3(parts, np), dim3(256), 0, b->stream, b->rtmp,
                       (int)b->nplane_r, b->nz, b->dp, f, w);
  }
  HIPC(hipGetLastError());
  return 0;
}

void rotate_history(qgk_bucket* b) {
  // after a step: new dqhdt_p lives in the old dqhdt_pp buffer
  int oldP = b->iP, oldPP = b->iPP;
  b->iP = oldPP;
  b->iPP = oldP;
}

// one fused step (row kernel + column kernel); qh must be primed
int fused_step(qgk_bucket* b, StepCoeffs cf) {
  hipEvent_t e0 = nullptr;
  int rc = t_begin(b, &e0);
  if (rc) return rc;
  rc = launch_row_flux(b);
  if (rc) return rc;
  rc = t_end(b, e0, 0);
  if (rc) return rc;
  rc = t_begin(b, &e0);
  if (rc) return rc;
  rc = launch_col_step(b, cf);
  if (rc) return rc;
  rc = t_end(b, e0, 1);
  if (rc) return rc;
  rotate_history(b);
  std::swap(b->field_c[QGK_QH], b->qh_alt);
  return 0;
}

int check_field(qgk_bucket* b, int f) {
  if (f < 0 || f >= QGK_NFIELDS) return fail(QGK_E_ARG, "bad field id %d", f);
  if (f == QGK_DQHDT || f == QGK_DQHDT_P || f == QGK_DQHDT_PP || f == QGK_QH) return 0;
  if (is_real_field(f) ? b->field_r[f] == nullptr : b->field_c[f] == nullptr)
    return fail(QGK_E_STATE, "field %d is not allocated for this handle", f);
  return 0;
}

cd* spectral_ptr(qgk_bucket* b, int f) {
  switch (f) {
    case QGK_DQHDT: return D(b);
    case QGK_DQHDT_P: return Pp(b);
    case QGK_DQHDT_PP: return PP(b);
    default: return b->field_c[f];
  }
}

int refresh_q(qgk_bucket* b) {
  if (!b->q_stale) return 0;
  int rc = c2r2d(b, b->field_c[QGK_QH], b->field_r[QGK_Q], b->M * b->nz);
  if (rc) return rc;
  b->q_stale = false;
  return 0;
}

// Host tables are laid out with nk columns; device tables with ldk.  Expand
// `rows` rows of nk values into rows of ldk values (zero padding).
std::vector<double> pad_rows(const double* host, size_t rows, int nk, int ldk) {
  std::vector<double> out(rows * ldk, 0.0);
  for (size_t r = 0; r < rows; ++r)
    std::memcpy(out.data() + r * ldk, host + r * nk, nk * sizeof(double));
  return out;
}

// rows of an (m members x) table in host layout for constant `which`
size_t const_rows(const qgk_bucket* b, int which) {
  switch (which) {
    case QGK_KK: return 1;
    case QGK_K2L2:
    case QGK_FILTR: return b->nl;
    case QGK_A: return (size_t)b->nz * b->nz * b->nl;
    default: return 0;  // LL, UBG, QY: not padded
  }
}

// spectral copies between a dense host layout [rows][nk] and the padded device one
hipError_t copy_spectral(qgk_bucket* b, void* dst, const void* src, size_t rows,
                         hipMemcpyKind kind) {
  const size_t dense = (size_t)b->nk * sizeof(cd), padded = (size_t)b->ldk * sizeof(cd);
  const size_t dpitch = kind == hipMemcpyHostToDevice || kind == hipMemcpyDeviceToDevice ?
                        padded : dense;
  const size_t spitch = kind == hipMemcpyDeviceToHost || kind == hipMemcpyDeviceToDevice ?
                        padded : dense;
  return hipMemcpy2DAsync(dst, dpitch, src, spitch, dense, rows, kind, b->stream);
}

int need_full(qgk_bucket* b, const char* what) {
  if (!(b->flags & QGK_FULL_STATE))
    return fail(QGK_E_STATE, "%s needs a handle created with QGK_FULL_STATE", what);
  return 0;
}

}  // namespace

// ==========================================================================
// C-ABI
// ==========================================================================
extern "C" {

const char* qgk_last_error(void) { return g_err.c_str(); }

int qgk_create(int nmembers, int nz, int ny, int nx, int flags, qgk_bucket** out) {
  if (!out) return fail(QGK_E_ARG, "out is NULL");
  *out = nullptr;
  if (nmembers < 1 || nz < 1) return fail(QGK_E_ARG, "nmembers and nz must be >= 1");
  if (!size_supported(nx) || !size_supported(ny))
    return fail(QGK_E_ARG,
                "nx=%d, ny=%d: supported sizes are 2^a in [8, 1024] and 3*2^a in [24, 768]",
                nx, ny);
  int ndev = 0;
  hipError_t he = hipGetDeviceCount(&ndev);
  if (he != hipSuccess || ndev == 0)
    return fail(QGK_E_HIP, "no HIP device available (%s)", hipGetErrorString(he));
  qgk_bucket* b = new qgk_bucket();
  b->M = nmembers;
  b->nz = nz;
  b->ny = ny;
  b->nx = nx;
  b->nl = ny;
  b->nk = nx / 2 + 1;
  b->ldk = ldk_of(nx);
  b->flags = flags;
  const long nmax = nx > ny ? nx : ny;
  if (nmembers == 1 || (long)nmembers * nmax * nmax <= (1L << 19)) {
    // Small buckets: the fused launches are few workgroups and bound by the
    // latency of their serial transform passes.  Radix 4 (twice the lanes
    // per transform) and, with two layers, the layer-split column kernel
    // measured fastest up to members x n^2 = 2^19: per step 20-50 us against
    // 31-62 us for radix 8 and the register kernel at n = 64 with 1 ... 128
    // members, 41-55 against 52-64 us at n = 256 with 1 ... 8; beyond that
    // (n = 128 with 64 members, n = 256 with 16) the split kernel loses.
    // Larger buckets keep radix 8 and the register kernel (the autotuner
    // may change the radix).
    b->erow = 4;
    b->ecol = 4;
    b->csplit = nz == 2;
  }
  b->nplane_c = (size_t)b->nl * b->ldk;
  b->nplane_r = (size_t)ny * nx;
  int rc = 0;
  auto cleanup = [&](int code) {
    qgk_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(QGK_E_HIP, "hipStreamCreate failed"));
  {  // set-up of the allocation-free entry point's scratch address on this device
    unsigned long long* w = nullptr;
    if ((rc = scratch_address(&w))) return cleanup(rc);
  }
  const size_t nc = (size_t)nmembers * nz * b->nplane_c * sizeof(cd);
  const size_t nr = (size_t)nmembers * nz * b->nplane_r * sizeof(double);
  auto ac = [&](cd** p) { return dev_alloc(b, (void**)p, nc); };
  auto ar = [&](double** p) { return dev_alloc(b, (void**)p, nr); };
  if ((rc = ac(&b->field_c[QGK_QH])) || (rc = ac(&b->qh_alt)) || (rc = ac(&b->hist[0])) || (rc = ac(&b->hist[1])) ||
      (rc = ac(&b->hist[2])) || (rc = ac(&b->Qm)) || (rc = ac(&b->Um)) || (rc = ac(&b->Vm)) ||
      (rc = ac(&b->UQm)) || (rc = ac(&b->VQm)) || (rc = ac(&b->ctmp)) || (rc = ar(&b->rtmp)) ||
      (rc = ar(&b->field_r[QGK_Q])))
    return cleanup(rc);
  if (nz >= 3 && nz <= kTileNzMax && (rc = ac(&b->phm))) return cleanup(rc);
  if (flags & QGK_FULL_STATE) {
    const int cf[] = {QGK_PH, QGK_UH, QGK_VH, QGK_UQH, QGK_VQH};
    const int rf[] = {QGK_U, QGK_V, QGK_UQ, QGK_VQ};
    for (int f : cf)
      if ((rc = ac(&b->field_c[f]))) return cleanup(rc);
    for (int f : rf)
      if ((rc = ar(&b->field_r[f]))) return cleanup(rc);
  }
  if (flags & QGK_HAS_Q_PARAM) {
    if ((rc = ar(&b->field_r[QGK_DQ])) || (rc = ac(&b->field_c[QGK_DQH]))) return cleanup(rc);
  }
  if (flags & QGK_HAS_UV_PARAM) {
    if ((rc = ar(&b->field_r[QGK_DU])) || (rc = ar(&b->field_r[QGK_DV])) ||
        (rc = ac(&b->field_c[QGK_DUH])) || (rc = ac(&b->field_c[QGK_DVH])))
      return cleanup(rc);
  }
  // twiddles
  {
    auto twx = make_twiddles(nx), twy = make_twiddles(ny);
    if ((rc = dev_alloc(b, (void**)&b->tw_x, nx * sizeof(cd))) ||
        (rc = dev_alloc(b, (void**)&b->tw_y, ny * sizeof(cd))))
      return cleanup(rc);
    if (hipMemcpyAsync(b->tw_x, twx.data(), nx * sizeof(cd), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipMemcpyAsync(b->tw_y, twy.data(), ny * sizeof(cd), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return cleanup(fail(QGK_E_HIP, "twiddle upload failed"));
  }
  // default constants: zeros, filtr = 1 (shared tables)
  b->own_tab.assign(QGK_NCONSTS, std::vector<double*>(nmembers, nullptr));
  for (int w = 0; w < QGK_NCONSTS; ++w) {
    const size_t n = const_len(b, w);
    if ((rc = dev_alloc(b, (void**)&b->shared_tab[w], n * sizeof(double)))) return cleanup(rc);
    if (w == QGK_FILTR) {
      // same stream as the zero fill above, then wait: `ones` is pageable and local
      std::vector<double> ones(n, 1.0);
      if (hipMemcpyAsync(b->shared_tab[w], ones.data(), n * sizeof(double),
                         hipMemcpyHostToDevice, b->stream) != hipSuccess ||
          hipStreamSynchronize(b->stream) != hipSuccess)
        return cleanup(fail(QGK_E_HIP, "filter init failed"));
    }
  }
  b->hp.assign(nmembers, MemberParams{});
  for (auto& p : b->hp) {
    for (int w = 0; w < QGK_NCONSTS; ++w) param_slot(p, w) = b->shared_tab[w];
    p.rek = 0.0;
  }
  if ((rc = dev_alloc(b, (void**)&b->dp, sizeof(MemberParams) * nmembers))) return cleanup(rc);
  if ((rc = dev_alloc(b, (void**)&b->small, 256 * sizeof(double)))) return cleanup(rc);
  b->params_dirty = true;
  if (hipStreamSynchronize(b->stream) != hipSuccess)
    return cleanup(fail(QGK_E_HIP, "stream sync failed"));
  *out = b;
  return 0;
}

int qgk_destroy(qgk_bucket* b) {
  if (!b) return 0;
  if (b->stream) hipStreamSynchronize(b->stream);
  for (auto& g : b->gexec)
    if (g) hipGraphExecDestroy(g);
  if (b->graph.exec) hipGraphExecDestroy(b->graph.exec);
  if (b->ugexec) hipGraphExecDestroy(b->ugexec);
  for (auto& p : b->pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (void* p : b->allocs) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int qgk_set_const(qgk_bucket* b, int which, int member, const double* host) {
  if (!b || !host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "setting a constant")) return rc0;
  if (which < 0 || which >= QGK_NCONSTS) return fail(QGK_E_ARG, "bad constant id %d", which);
  if (member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  const size_t bytes = const_len(b, which) * sizeof(double);
  double* dst;
  if (member < 0) {
    dst = b->shared_tab[which];
    for (int m = 0; m < b->M; ++m) param_slot(b->hp[m], which) = dst;
  } else {
    double*& own = b->own_tab[which][member];
    if (!own) {
      int rc = dev_alloc(b, (void**)&own, bytes);
      if (rc) return rc;
    }
    dst = own;
    param_slot(b->hp[member], which) = dst;
  }
  if (which == QGK_A) {
    b->gvalid = false;
    for (int m = (member < 0 ? 0 : member); m < (member < 0 ? b->M : member + 1); ++m)
      b->hp[m].tri = nullptr;  // a new table is used as given
  }
  if (which == QGK_LL)
    for (int m = (member < 0 ? 0 : member); m < (member < 0 ? b->M : member + 1); ++m)
      b->hp[m].dl = b->nl > 1 ? host[1] : 0.0;
  b->gen++;
  b->params_dirty = true;
  const size_t rows = const_rows(b, which);
  if (rows) {
    std::vector<double> tab = pad_rows(host, rows, b->nk, b->ldk);
    HIPC(hipMemcpyAsync(dst, tab.data(), bytes, hipMemcpyHostToDevice, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
  } else {
    HIPC(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
  }
  b->primed = false;
  return 0;
}

int qgk_set_const_all(qgk_bucket* b, int which, const double* host) {
  if (!b || !host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "setting a constant")) return rc0;
  if (which < 0 || which >= QGK_NCONSTS) return fail(QGK_E_ARG, "bad constant id %d", which);
  const size_t n = const_len(b, which);
  double* tab = nullptr;
  int rc = dev_alloc(b, (void**)&tab, n * b->M * sizeof(double));
  if (rc) return rc;
  const size_t rows = const_rows(b, which);
  std::vector<double> buf;
  const double* src = host;
  if (rows) {
    buf = pad_rows(host, rows * b->M, b->nk, b->ldk);
    src = buf.data();
  }
  HIPC(hipMemcpyAsync(tab, src, n * b->M * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  for (int m = 0; m < b->M; ++m) {
    param_slot(b->hp[m], which) = tab + (size_t)m * n;
    if (which == QGK_A) b->hp[m].tri = nullptr;
    if (which == QGK_LL) b->hp[m].dl = b->nl > 1 ? host[(size_t)m * n + 1] : 0.0;
  }
  b->params_dirty = true;
  b->primed = false;
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_set_inversion_2layer(qgk_bucket* b, int member, double F1, double F2) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "setting the inversion")) return rc0;
  if (b->nz != 2) return fail(QGK_E_ARG, "closed-form inversion needs nz == 2");
  if (member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  for (int m = (member < 0 ? 0 : member); m < (member < 0 ? b->M : member + 1); ++m) {
    b->hp[m].a = nullptr;
    b->hp[m].F1 = F1;
    b->hp[m].F2 = F2;
  }
  b->params_dirty = true;
  b->primed = false;
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_set_inversion_tridiagonal(qgk_bucket* b, int member, const double* bands) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "setting the inversion")) return rc0;
  if (b->nz < 3) return fail(QGK_E_ARG, "the tridiagonal solve is for nz >= 3 (nz=%d)", b->nz);
  if (member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  const int m0 = member < 0 ? 0 : member, m1 = member < 0 ? b->M : member + 1;
  if (bands) {
    for (int i = 0; i < b->nz; ++i) {
      const double lo = bands[i], dg = bands[b->nz + i], hi = bands[2 * b->nz + i];
      if ((i == 0 && lo != 0.0) || (i == b->nz - 1 && hi != 0.0))
        return fail(QGK_E_ARG, "bands[0][0] and bands[2][nz-1] must be 0");
      if (!(dg <= 0.0) || lo < 0.0 || hi < 0.0 || -dg < lo + hi)
        return fail(QGK_E_ARG, "layer %d: the solve needs a diagonally dominant S with a "
                               "non-positive diagonal and non-negative off-diagonals", i);
    }
    const size_t per = (size_t)3 * b->nz;
    if (!b->tri_tab) {
      int rc = dev_alloc(b, (void**)&b->tri_tab, per * b->M * sizeof(double));
      if (rc) return rc;
    }
    for (int m = m0; m < m1; ++m)
      HIPC(hipMemcpyAsync(b->tri_tab + m * per, bands, per * sizeof(double),
                          hipMemcpyHostToDevice, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
  }
  for (int m = m0; m < m1; ++m) b->hp[m].tri = bands ? b->tri_tab + m * (size_t)3 * b->nz : nullptr;
  b->params_dirty = true;
  b->primed = false;
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_get_inversion_tridiagonal(qgk_bucket* b, int member, int* on) {
  if (!b || !on) return fail(QGK_E_ARG, "NULL argument");
  if (member < 0 || member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  *on = b->hp[member].tri != nullptr;
  return 0;
}

int qgk_set_rek(qgk_bucket* b, int member, double rek) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "setting rek")) return rc0;
  if (member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  for (int m = (member < 0 ? 0 : member); m < (member < 0 ? b->M : member + 1); ++m)
    b->hp[m].rek = rek;
  b->params_dirty = true;
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_put(qgk_bucket* b, int field, const void* host) {
  if (!b || !host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_put")) return rc0;
  int rc = check_field(b, field);
  if (rc) return rc;
  const size_t n = (size_t)b->M * b->nz;
  if (is_real_field(field)) {
    HIPC(hipMemcpyAsync(b->field_r[field], host, n * b->nplane_r * sizeof(double),
                        hipMemcpyHostToDevice, b->stream));
  } else {
    HIPC(copy_spectral(b, spectral_ptr(b, field), host, n * b->nl, hipMemcpyHostToDevice));
  }
  HIPC(hipStreamSynchronize(b->stream));
  if (field == QGK_QH) b->primed = false;
  if (field == QGK_Q) b->q_stale = false;
  return 0;
}

int qgk_get(qgk_bucket* b, int field, void* host) {
  if (!b || !host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_get")) return rc0;
  int rc = check_field(b, field);
  if (rc) return rc;
  if (field == QGK_Q && (rc = refresh_q(b))) return rc;
  const size_t n = (size_t)b->M * b->nz;
  if (is_real_field(field)) {
    HIPC(hipMemcpyAsync(host, b->field_r[field], n * b->nplane_r * sizeof(double),
                        hipMemcpyDeviceToHost, b->stream));
  } else {
    HIPC(copy_spectral(b, host, spectral_ptr(b, field), n * b->nl, hipMemcpyDeviceToHost));
  }
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

// End of an entry point that only enqueues device work: wait for it, unless
// the stream is being captured into a graph (then nothing runs yet).
// The four phases of a plain step (qgk_invert, qgk_advect, qgk_friction,
// qgk_timestep) do not wait: they take no host data, every later kernel is
// ordered after them on the bucket's stream, and every entry point that hands
// data to or from the host waits for the stream first, so a loop of
// _step_forward calls queues its kernels back to back instead of waiting for
// each phase (enqueue-only phases; launch errors are still returned here,
// errors during execution by the next waiting call).
hipError_t stream_done(qgk_bucket* b) {
  return b->capturing ? hipSuccess : hipStreamSynchronize(b->stream);
}

int qgk_invert(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  int rc = need_full(b, "_invert");
  if (rc || (rc = upload_params(b))) return rc;
  const size_t pts = (size_t)b->M * b->nplane_c;
  hipLaunchKernelGGL(k_invert, dim3(grid_1d(pts)), dim3(256), 0, b->stream, b->field_c[QGK_QH],
                     b->field_c[QGK_PH], b->field_c[QGK_UH], b->field_c[QGK_VH], b->dp, b->nz,
                     b->nl, b->ldk, b->M);
  HIPC(hipGetLastError());
  const int np = b->M * b->nz;
  {
    const cd* in[2] = {b->field_c[QGK_UH], b->field_c[QGK_VH]};
    double* out[2] = {b->field_r[QGK_U], b->field_r[QGK_V]};
    if ((rc = c2r2d_n(b, 2, in, out, np))) return rc;
  }
  HIPC(hipGetLastError());  // no wait: see enqueue-only phases above
  return 0;
}

int qgk_advect(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  int rc = need_full(b, "_do_advection");
  if (rc || (rc = upload_params(b)) || (rc = refresh_q(b))) return rc;
  const int np = b->M * b->nz;
  const size_t rpts = (size_t)np * b->nplane_r;
  hipLaunchKernelGGL(k_flux, dim3(grid_1d(rpts)), dim3(256), 0, b->stream, b->field_r[QGK_U],
                     b->field_r[QGK_V], b->field_r[QGK_Q], b->field_r[QGK_UQ],
                     b->field_r[QGK_VQ], b->dp, b->nz, b->nplane_r, b->M);
  HIPC(hipGetLastError());
  {
    const double* in[2] = {b->field_r[QGK_UQ], b->field_r[QGK_VQ]};
    cd* out[2] = {b->field_c[QGK_UQH], b->field_c[QGK_VQH]};
    if ((rc = r2c2d_n(b, 2, in, out, np))) return rc;
  }
  const size_t cpts = (size_t)np * b->nplane_c;
  hipLaunchKernelGGL(k_divergence, dim3(grid_1d(cpts)), dim3(256), 0, b->stream,
                     b->field_c[QGK_UQH], b->field_c[QGK_VQH], b->field_c[QGK_PH], D(b), b->dp,
                     b->nz, b->nl, b->ldk, b->M);
  HIPC(hipGetLastError());
  HIPC(hipGetLastError());  // no wait: see enqueue-only phases above
  return 0;
}

int qgk_friction(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  int rc = need_full(b, "_do_friction");
  if (rc || (rc = upload_params(b))) return rc;
  const size_t pts = (size_t)b->M * b->nplane_c;
  hipLaunchKernelGGL(k_friction, dim3(grid_1d(pts)), dim3(256), 0, b->stream,
                     b->field_c[QGK_PH], D(b), b->dp, b->nz, b->nl, b->ldk, b->M);
  HIPC(hipGetLastError());
  HIPC(hipGetLastError());  // no wait: see enqueue-only phases above
  return 0;
}

int qgk_apply_q_param(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (!b->field_r[QGK_DQ]) return fail(QGK_E_STATE, "dq is not allocated (no q parameterization)");
  const int np = b->M * b->nz;
  int rc = r2c2d(b, b->field_r[QGK_DQ], b->field_c[QGK_DQH], np);
  if (rc) return rc;
  const size_t cpts = (size_t)np * b->nplane_c;
  hipLaunchKernelGGL(k_add, dim3(grid_1d(cpts)), dim3(256), 0, b->stream, D(b),
                     b->field_c[QGK_DQH], cpts);
  HIPC(hipGetLastError());
  HIPC(stream_done(b));
  return 0;
}

int qgk_add_q_param(qgk_bucket* b, const double* dq_host) {
  if (!b || !dq_host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "a host dq")) return rc0;
  if (!b->field_r[QGK_DQ]) return fail(QGK_E_STATE, "dq is not allocated (no q parameterization)");
  const int np = b->M * b->nz;
  HIPC(hipMemcpyAsync(b->field_r[QGK_DQ], dq_host, (size_t)np * b->nplane_r * sizeof(double),
                      hipMemcpyHostToDevice, b->stream));
  return qgk_apply_q_param(b);
}

int qgk_apply_uv_param(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (!b->field_r[QGK_DU])
    return fail(QGK_E_STATE, "du/dv are not allocated (no uv parameterization)");
  int rc = upload_params(b);
  if (rc) return rc;
  const int np = b->M * b->nz;
  {
    const double* in[2] = {b->field_r[QGK_DU], b->field_r[QGK_DV]};
    cd* out[2] = {b->field_c[QGK_DUH], b->field_c[QGK_DVH]};
    if ((rc = r2c2d_n(b, 2, in, out, np))) return rc;
  }
  const size_t cpts = (size_t)np * b->nplane_c;
  hipLaunchKernelGGL(k_add_curl, dim3(grid_1d(cpts)), dim3(256), 0, b->stream, D(b),
                     b->field_c[QGK_DUH], b->field_c[QGK_DVH], b->dp, b->nz, b->nl, b->ldk, b->M);
  HIPC(hipGetLastError());
  HIPC(stream_done(b));
  return 0;
}

int qgk_add_uv_param(qgk_bucket* b, const double* du_host, const double* dv_host) {
  if (!b || !du_host || !dv_host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "a host du, dv")) return rc0;
  if (!b->field_r[QGK_DU])
    return fail(QGK_E_STATE, "du/dv are not allocated (no uv parameterization)");
  const int np = b->M * b->nz;
  const size_t bytes = (size_t)np * b->nplane_r * sizeof(double);
  HIPC(hipMemcpyAsync(b->field_r[QGK_DU], du_host, bytes, hipMemcpyHostToDevice, b->stream));
  HIPC(hipMemcpyAsync(b->field_r[QGK_DV], dv_host, bytes, hipMemcpyHostToDevice, b->stream));
  return qgk_apply_uv_param(b);
}

int qgk_timestep(qgk_bucket* b, double dt1, double dt2, double dt3) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  int rc = upload_params(b);
  if (rc) return rc;
  const int np = b->M * b->nz;
  const size_t cpts = (size_t)np * b->nplane_c;
  StepCoeffs c = {dt1, dt2, dt3};
  hipLaunchKernelGGL(k_ab_step, dim3(grid_1d(cpts)), dim3(256), 0, b->stream,
                     b->field_c[QGK_QH], D(b), Pp(b), PP(b), b->dp, c, b->nz, b->nl, b->ldk,
                     b->M);
  HIPC(hipGetLastError());
  rotate_history(b);
  b->primed = false;
  b->q_stale = false;
  if ((rc = c2r2d(b, b->field_c[QGK_QH], b->field_r[QGK_Q], np))) return rc;
  HIPC(hipGetLastError());  // no wait: see enqueue-only phases above
  return 0;
}

int qgk_capture_begin(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (b->capturing) return fail(QGK_E_STATE, "already capturing");
  int rc = need_full(b, "graph capture");
  if (rc || (rc = upload_params(b)) || (rc = refresh_q(b))) return rc;
  HIPC(hipStreamSynchronize(b->stream));
  HIPC(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  b->capturing = true;
  b->ug_phase = b->iP;
  b->ug_gen = b->gen;
  b->ug_qh = b->field_c[QGK_QH];
  b->ug_fused = false;
  return 0;
}

int qgk_capture_end(qgk_bucket* b, int keep) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (!b->capturing) return fail(QGK_E_STATE, "not capturing");
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  b->capturing = false;
  if (b->ugexec) {
    hipGraphExecDestroy(b->ugexec);
    b->ugexec = nullptr;
  }
  int rc = 0;
  if (e != hipSuccess || !g) {
    rc = fail(QGK_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  } else if (keep) {
    if (b->iP != b->ug_phase)
      rc = fail(QGK_E_STATE, "a step graph must cover whole cycles of the history rotation "
                             "(a multiple of 2 steps)");
    else if (b->gen != b->ug_gen)
      rc = fail(QGK_E_STATE, "constants changed during graph capture");
    else if (hipGraphInstantiate(&b->ugexec, g, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(QGK_E_HIP, "graph instantiation failed");
  }
  if (g) hipGraphDestroy(g);
  if (rc) b->ugexec = nullptr;
  return rc;
}

int qgk_graph_replay(qgk_bucket* b, int n) {
  if (!b || n < 0) return fail(QGK_E_ARG, "bad arguments");
  if (b->capturing) return fail(QGK_E_STATE, "capture in progress");
  if (!b->ugexec) return fail(QGK_E_STATE, "no captured graph");
  if (b->iP != b->ug_phase || b->gen != b->ug_gen || b->field_c[QGK_QH] != b->ug_qh)
    return fail(QGK_E_STATE,
                "the captured graph is stale (history phase, qh buffer or constants changed)");
  int rc = upload_params(b);
  if (rc) return rc;
  for (int i = 0; i < n; ++i) HIPC(hipGraphLaunch(b->ugexec, b->stream));
  // the graph ends with a time step, as qgk_timestep leaves the bucket, or
  // with a fused closed step, as qgk_fused_col_forcing leaves it
  if (n > 0) {
    b->primed = b->ug_fused;
    b->q_stale = b->ug_fused;
  }
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_q_to_qh(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_q_to_qh")) return rc0;
  int rc = r2c2d(b, b->field_r[QGK_Q], b->field_c[QGK_QH], b->M * b->nz);
  if (rc) return rc;
  b->primed = false;
  b->q_stale = false;
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_qh_to_q(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_qh_to_q")) return rc0;
  int rc = c2r2d(b, b->field_c[QGK_QH], b->field_r[QGK_Q], b->M * b->nz);
  if (rc) return rc;
  b->q_stale = false;
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_fft(qgk_bucket* b, int nplanes, const double* in_real, double* out_complex) {
  if (!b || !in_real || !out_complex) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_fft")) return rc0;
  if (nplanes < 1 || nplanes > b->M * b->nz)
    return fail(QGK_E_ARG, "nplanes must be in [1, %d]", b->M * b->nz);
  HIPC(hipMemcpyAsync(b->rtmp, in_real, (size_t)nplanes * b->nplane_r * sizeof(double),
                      hipMemcpyHostToDevice, b->stream));
  // keep the generic transform off every state buffer: UQm is free between steps
  // only in the fused path, so use the dedicated scratch pair (rtmp -> VQm is
  // not safe either); the row pass writes its half spectra into ctmp.
  int rc = launch_row_r2c(b, b->rtmp, b->ctmp, nplanes);
  if (rc) return rc;
  if ((rc = launch_col_fft<-1>(b, b->ctmp, b->ctmp, nplanes))) return rc;
  HIPC(copy_spectral(b, out_complex, b->ctmp, (size_t)nplanes * b->nl, hipMemcpyDeviceToHost));
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_ifft(qgk_bucket* b, int nplanes, const double* in_complex, double* out_real) {
  if (!b || !in_complex || !out_real) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_ifft")) return rc0;
  if (nplanes < 1 || nplanes > b->M * b->nz)
    return fail(QGK_E_ARG, "nplanes must be in [1, %d]", b->M * b->nz);
  HIPC(copy_spectral(b, b->ctmp, in_complex, (size_t)nplanes * b->nl, hipMemcpyHostToDevice));
  int rc = launch_col_fft<+1>(b, b->ctmp, b->ctmp, nplanes);
  if (rc) return rc;
  if ((rc = launch_row_c2r(b, b->ctmp, b->rtmp, nplanes))) return rc;
  HIPC(hipMemcpyAsync(out_real, b->rtmp, (size_t)nplanes * b->nplane_r * sizeof(double),
                      hipMemcpyDeviceToHost, b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_timing_enable(qgk_bucket* b, int enable) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_timing_enable")) return rc0;
  int rc = t_collect(b);
  if (rc) return rc;
  b->timing = enable != 0;
  b->ms_row = b->ms_col = 0;
  b->n_row = b->n_col = 0;
  return 0;
}

int qgk_timing(qgk_bucket* b, double* ms_row, long* n_row, double* ms_col, long* n_col) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_timing")) return rc0;
  int rc = t_collect(b);
  if (rc) return rc;
  if (ms_row) *ms_row = b->ms_row;
  if (n_row) *n_row = b->n_row;
  if (ms_col) *ms_col = b->ms_col;
  if (n_col) *n_col = b->n_col;
  return 0;
}

int qgk_sync(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_sync")) return rc0;
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

namespace {

bool graph_matches(const qgk_bucket::GraphCache& g, qgk_bucket** bs, int nb, StepCoeffs cf) {
  if (!g.exec || (int)g.bs.size() != nb) return false;
  if (g.cf.dt1 != cf.dt1 || g.cf.dt2 != cf.dt2 || g.cf.dt3 != cf.dt3) return false;
  for (int i = 0; i < nb; ++i)
    if (g.bs[i] != bs[i] || g.gens[i] != bs[i]->gen || g.qh0[i] != bs[i]->field_c[QGK_QH] ||
        g.iP0[i] != bs[i]->iP)
      return false;
  return true;
}

// Capture two fused steps of every bucket on bs[0]'s stream.  The history and
// qh buffers rotate with period two, so after the capture the host-side
// rotation state is back where it started and the graph can be replayed.
int capture_two_steps(qgk_bucket** bs, int nb, StepCoeffs cf) {
  qgk_bucket::GraphCache& g = bs[0]->graph;
  if (g.exec) {
    hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
  }
  hipStream_t cs = bs[0]->stream;
  for (int i = 0; i < nb; ++i) HIPC(hipStreamSynchronize(bs[i]->stream));
  HIPC(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  int rc = 0;
  for (int s = 0; s < 2 && rc == 0; ++s)
    for (int i = 0; i < nb && rc == 0; ++i) {
      bs[i]->lst = cs;
      if ((rc = launch_row_flux(bs[i])) == 0 && (rc = launch_col_step(bs[i], cf)) == 0) {
        rotate_history(bs[i]);
        std::swap(bs[i]->field_c[QGK_QH], bs[i]->qh_alt);
      }
      bs[i]->lst = nullptr;
    }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc) {
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return fail(QGK_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) {
    g.exec = nullptr;
    return fail(QGK_E_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
  }
  g.bs.assign(bs, bs + nb);
  g.gens.resize(nb);
  g.qh0.resize(nb);
  g.iP0.resize(nb);
  for (int i = 0; i < nb; ++i) {
    g.gens[i] = bs[i]->gen;
    g.qh0[i] = bs[i]->field_c[QGK_QH];
    g.iP0[i] = bs[i]->iP;
  }
  g.cf = cf;
  return 0;
}

}  // namespace

int qgk_step_buckets(qgk_bucket** bs, int nb, int nsteps, const double* coeffs, int use_graph) {
  if (!bs || nb < 1 || nsteps < 0 || (nsteps > 0 && !coeffs))
    return fail(QGK_E_ARG, "bad arguments");
  bool timing = false;
  for (int i = 0; i < nb; ++i) {
    qgk_bucket* b = bs[i];
    if (!b) return fail(QGK_E_ARG, "NULL bucket");
    if (int rc0 = refuse_in_capture(b, "a fused step")) return rc0;
    int rc = upload_params(b);
    if (rc) return rc;
    if (!b->primed) {
      if ((rc = launch_col_prologue(b))) return rc;
      b->primed = true;
    }
    timing = timing || b->timing;
  }
  // the trailing run of steps with identical AB weights may replay a graph
  int first_const = nsteps;
  while (first_const > 0 && coeffs[3 * (first_const - 1)] == coeffs[3 * (nsteps - 1)] &&
         coeffs[3 * (first_const - 1) + 1] == coeffs[3 * (nsteps - 1) + 1] &&
         coeffs[3 * (first_const - 1) + 2] == coeffs[3 * (nsteps - 1) + 2])
    --first_const;
  const bool graph = use_graph && !timing && nsteps - first_const >= 2;
  const int eager_end = graph ? first_const : nsteps;
  for (int s = 0; s < eager_end; ++s) {
    StepCoeffs cf = {coeffs[3 * s], coeffs[3 * s + 1], coeffs[3 * s + 2]};
    for (int i = 0; i < nb; ++i) {
      int rc = fused_step(bs[i], cf);
      if (rc) return rc;
    }
  }
  if (graph) {
    const int s0 = first_const;
    StepCoeffs cf = {coeffs[3 * s0], coeffs[3 * s0 + 1], coeffs[3 * s0 + 2]};
    if (!graph_matches(bs[0]->graph, bs, nb, cf)) {
      int rc = capture_two_steps(bs, nb, cf);
      if (rc) return rc;
    }
    // the other buckets' pending work must precede the replays on bs[0]'s stream
    for (int i = 1; i < nb; ++i) HIPC(hipStreamSynchronize(bs[i]->stream));
    const int pairs = (nsteps - s0) / 2;
    for (int r = 0; r < pairs; ++r) HIPC(hipGraphLaunch(bs[0]->graph.exec, bs[0]->stream));
    if ((nsteps - s0) % 2) {
      HIPC(hipStreamSynchronize(bs[0]->stream));
      for (int i = 0; i < nb; ++i) {
        int rc = fused_step(bs[i], cf);
        if (rc) return rc;
      }
    }
  }
  for (int i = 0; i < nb; ++i) {
    // the fused step leaves the real q stale; it is refreshed when next needed
    if (nsteps > 0) bs[i]->q_stale = true;
    HIPC(hipStreamSynchronize(bs[i]->stream));
  }
  return 0;
}

int qgk_copy_probe(size_t nbytes, int iters, double* ms_per_copy) {
  if (nbytes < 16 || iters < 1 || !ms_per_copy) return fail(QGK_E_ARG, "bad arguments");
  const size_t n = nbytes / 16;
  double2 *a = nullptr, *c = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  if (hipMalloc((void**)&a, n * 16) != hipSuccess || hipMalloc((void**)&c, n * 16) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(QGK_E_HIP, "copy probe set-up failed");
  } else {
    hipMemsetAsync(a, 0, n * 16, st);
    double best = 0.0;
    const size_t nblk = (n + 255) / 256;
    if (nblk > 0x7fffffffu) {
      rc = fail(QGK_E_ARG, "copy probe size too large");
    } else {
      const dim3 grid((unsigned)nblk);
      hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, st, a, c, n);
      hipEventRecord(e0, st);
      for (int i = 0; i < iters; ++i)
        hipLaunchKernelGGL(k_stream_copy, grid, dim3(256), 0, st, a, c, n);
      hipEventRecord(e1, st);
      float ms = 0;
      if (hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(QGK_E_HIP, "copy probe timing failed");
      } else {
        best = ms / iters;
      }
    }
    *ms_per_copy = best;
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  if (a) hipFree(a);
  if (c) hipFree(c);
  return rc;
}

int qgk_pattern_probe(int ny, int ldk, int ncols, int members, int nz, int iters,
                      double* ms_rows, double* ms_cols) {
  if (ny < 32 || ny % 32 || ldk < 8 || ldk % 8 || members < 1 || nz < 1 || iters < 1 ||
      !ms_rows || !ms_cols)
    return fail(QGK_E_ARG, "pattern probe: bad arguments");
  if (ncols < 8 || ncols > ldk || ncols % 8) return fail(QGK_E_ARG, "pattern probe: bad ncols");
  const size_t n = (size_t)members * nz * ny * ldk;
  const size_t nrows = (size_t)members * nz * ny;
  if ((n + 255) / 256 > 0x7fffffffu || members > 65535)
    return fail(QGK_E_ARG, "pattern probe: size too large");
  cd* f[9] = {};
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  for (int i = 0; ok && i < 9; ++i)
    ok = hipMalloc((void**)&f[i], n * sizeof(cd)) == hipSuccess &&
         hipMemsetAsync(f[i], 0, n * sizeof(cd), st) == hipSuccess;
  if (!ok) rc = fail(QGK_E_HIP, "pattern probe set-up failed");
  auto timed = [&](auto launch, double* out) {
    if (rc) return;
    launch();
    hipEventRecord(e0, st);
    for (int i = 0; i < iters; ++i) launch();
    hipEventRecord(e1, st);
    float ms = 0;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = fail(QGK_E_HIP, "pattern probe timing failed");
    else
      *out = ms / iters;
  };
  timed([&] {
    hipLaunchKernelGGL(k_probe_rows, dim3((unsigned)((nrows * ncols + 255) / 256)), dim3(256),
                       0, st, f[0], f[1], f[2], f[3], f[4], nrows, ncols, ldk);
  }, ms_rows);
  timed([&] {
    hipLaunchKernelGGL(k_probe_cols, dim3(ncols / 8, members), dim3(256), 0, st, f[0], f[1], f[2],
                       f[3], f[4], f[5], f[6], f[7], f[8], ny, ldk, nz);
  }, ms_cols);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  for (int i = 0; i < 9; ++i)
    if (f[i]) hipFree(f[i]);
  return rc;
}

int qgk_set_variant(qgk_bucket* b, int erow, int ecol) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_set_variant")) return rc0;
  auto okr = [](int e) { return e == 4 || e == 8 || e == 16; };
  if (!okr(erow) || !okr(ecol)) return fail(QGK_E_ARG, "radix must be 4, 8 or 16");
  b->erow = erow;
  b->ecol = ecol;
  b->primed = false;  // the x-mixed exchange format may change with the column kernel
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_set_layer_split(qgk_bucket* b, int on) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_set_layer_split")) return rc0;
  if (on && b->nz != 2) return fail(QGK_E_ARG, "the layer-split column kernel needs nz == 2");
  b->csplit = on != 0;
  b->primed = false;  // the x-mixed exchange format may change with the column kernel
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_get_layer_split(qgk_bucket* b, int* on) {
  if (!b || !on) return fail(QGK_E_ARG, "NULL argument");
  *on = b->csplit ? 1 : 0;
  return 0;
}

int qgk_set_packed_exchange(qgk_bucket* b, int on) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_set_packed_exchange")) return rc0;
  b->pk_off = on == 0;
  b->primed = false;  // the x-mixed exchange format changes
  b->gvalid = false;
  b->gen++;
  return 0;
}

int qgk_exchange_columns(qgk_bucket* b, int* ncols) {
  if (!b || !ncols) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_exchange_columns")) return rc0;
  if (int rc = upload_params(b)) return rc;  // settles the standard-grid check
  *ncols = packed_exchange(b) ? b->nk - 1 : b->nk;
  return 0;
}

int qgk_get_variant(qgk_bucket* b, int* erow, int* ecol) {
  if (!b || !erow || !ecol) return fail(QGK_E_ARG, "NULL argument");
  *erow = b->erow;
  *ecol = b->ecol;
  return 0;
}

int qgk_autotune(qgk_bucket* b, int reps, int* erow, int* ecol, double* ms_row3,
                 double* ms_col3) {
  if (!b || reps < 1 || !erow || !ecol) return fail(QGK_E_ARG, "bad arguments");
  if (int rc0 = refuse_in_capture(b, "qgk_autotune")) return rc0;
  int rc = upload_params(b);
  if (rc) return rc;
  const size_t bytes = (size_t)b->M * b->nz * b->nplane_c * sizeof(cd);
  cd* save = nullptr;
  HIPC(hipMalloc((void**)&save, 3 * bytes));
  HIPC(hipMemcpyAsync(save, b->field_c[QGK_QH], bytes, hipMemcpyDeviceToDevice, b->stream));
  HIPC(hipMemcpyAsync(save + bytes / sizeof(cd), Pp(b), bytes, hipMemcpyDeviceToDevice,
                      b->stream));
  HIPC(hipMemcpyAsync(save + 2 * (bytes / sizeof(cd)), PP(b), bytes, hipMemcpyDeviceToDevice,
                      b->stream));
  const int cand[3] = {4, 8, 16};
  double best_r = 1e30, best_c = 1e30;
  int br = b->erow, bc = b->ecol;
  const StepCoeffs zero = {0.0, 0.0, 0.0};  // qh <- filtr qh: timing only
  hipEvent_t e0, e1;
  HIPC(hipEventCreate(&e0));
  HIPC(hipEventCreate(&e1));
  const int keep_r = b->erow, keep_c = b->ecol;
  for (int i = 0; i < 3 && rc == 0; ++i) {
    const int e = cand[i];
    const bool valid_r = e <= b->nx && !(i > 0 && cand[i - 1] >= b->nx);
    const bool valid_c = e <= b->ny && !(i > 0 && cand[i - 1] >= b->ny);
    float ms = 0;
    // row kernel
    b->erow = e;
    if (valid_r) {
      if ((rc = launch_row_flux(b))) break;
      HIPC(hipEventRecord(e0, b->stream));
      for (int r = 0; r < reps && rc == 0; ++r) rc = launch_row_flux(b);
      HIPC(hipEventRecord(e1, b->stream));
      HIPC(hipEventSynchronize(e1));
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      if (ms_row3) ms_row3[i] = ms / reps;
      if (ms / reps < best_r) {
        best_r = ms / reps;
        br = e;
      }
    } else if (ms_row3) {
      ms_row3[i] = -1.0;
    }
    // column kernel (writes qh_alt and dqhdt_pp; restored below)
    b->ecol = e;
    if (valid_c) {
      if ((rc = launch_col_step(b, zero))) break;
      HIPC(hipEventRecord(e0, b->stream));
      for (int r = 0; r < reps && rc == 0; ++r) rc = launch_col_step(b, zero);
      HIPC(hipEventRecord(e1, b->stream));
      HIPC(hipEventSynchronize(e1));
      HIPC(hipEventElapsedTime(&ms, e0, e1));
      if (ms_col3) ms_col3[i] = ms / reps;
      if (ms / reps < best_c) {
        best_c = ms / reps;
        bc = e;
      }
    } else if (ms_col3) {
      ms_col3[i] = -1.0;
    }
  }
  b->erow = keep_r;
  b->ecol = keep_c;
  HIPC(hipMemcpyAsync(b->field_c[QGK_QH], save, bytes, hipMemcpyDeviceToDevice, b->stream));
  HIPC(hipMemcpyAsync(Pp(b), save + bytes / sizeof(cd), bytes, hipMemcpyDeviceToDevice,
                      b->stream));
  HIPC(hipMemcpyAsync(PP(b), save + 2 * (bytes / sizeof(cd)), bytes, hipMemcpyDeviceToDevice,
                      b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  hipFree(save);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc) return rc;
  b->erow = br;
  b->ecol = bc;
  b->primed = false;  // Qm/Um/Vm were overwritten by the trials
  b->gvalid = false;
  b->gen++;
  *erow = br;
  *ecol = bc;
  return 0;
}

int qgk_device_name(char* buf, int len) {
  if (!buf || len < 1) return fail(QGK_E_ARG, "bad buffer");
  hipDeviceProp_t prop;
  int dev = 0;
  HIPC(hipGetDevice(&dev));
  HIPC(hipGetDeviceProperties(&prop, dev));
  snprintf(buf, len, "%s:%s:%d", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return 0;
}

int qgk_put_device(qgk_bucket* b, int field, const void* dptr) {
  if (!b || !dptr) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_put_device")) return rc0;
  int rc = check_field(b, field);
  if (rc) return rc;
  const size_t n = (size_t)b->M * b->nz;
  if (is_real_field(field)) {
    HIPC(hipMemcpyAsync(b->field_r[field], dptr, n * b->nplane_r * sizeof(double),
                        hipMemcpyDeviceToDevice, b->stream));
  } else {
    HIPC(hipMemcpy2DAsync(spectral_ptr(b, field), b->ldk * sizeof(cd), dptr, b->nk * sizeof(cd),
                          b->nk * sizeof(cd), n * b->nl, hipMemcpyDeviceToDevice, b->stream));
  }
  HIPC(hipStreamSynchronize(b->stream));
  if (field == QGK_QH) b->primed = false;
  if (field == QGK_Q) b->q_stale = false;
  return 0;
}

int qgk_get_device(qgk_bucket* b, int field, void* dptr) {
  if (!b || !dptr) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_get_device")) return rc0;
  int rc = check_field(b, field);
  if (rc) return rc;
  if (field == QGK_Q && (rc = refresh_q(b))) return rc;
  const size_t n = (size_t)b->M * b->nz;
  if (is_real_field(field)) {
    HIPC(hipMemcpyAsync(dptr, b->field_r[field], n * b->nplane_r * sizeof(double),
                        hipMemcpyDeviceToDevice, b->stream));
  } else {
    HIPC(hipMemcpy2DAsync(dptr, b->nk * sizeof(cd), spectral_ptr(b, field), b->ldk * sizeof(cd),
                          b->nk * sizeof(cd), n * b->nl, hipMemcpyDeviceToDevice, b->stream));
  }
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_coarsen(qgk_bucket* hi, qgk_bucket* lo, int field_hi, int field_lo,
                const double* filter_lo) {
  if (!hi || !lo) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(hi, "qgk_coarsen")) return rc0;
  if (int rc0 = refuse_in_capture(lo, "qgk_coarsen")) return rc0;
  if (hi->M != lo->M || hi->nz != lo->nz || hi->nx != hi->ny || lo->nx != lo->ny ||
      lo->nx > hi->nx)
    return fail(QGK_E_ARG, "coarsening needs equal members/layers and square grids, lo <= hi");
  if (is_real_field(field_hi) || is_real_field(field_lo))
    return fail(QGK_E_ARG, "coarsening acts on spectral fields");
  int rc = check_field(hi, field_hi);
  if (rc || (rc = check_field(lo, field_lo))) return rc;
  double* G = nullptr;
  std::vector<double> gpad;
  if (filter_lo) {
    gpad = pad_rows(filter_lo, lo->nl, lo->nk, lo->ldk);
    HIPC(hipMalloc((void**)&G, lo->nplane_c * sizeof(double)));
    HIPC(hipMemcpyAsync(G, gpad.data(), lo->nplane_c * sizeof(double), hipMemcpyHostToDevice,
                        lo->stream));
  }
  HIPC(hipStreamSynchronize(hi->stream));
  const double s = ((double)lo->nx * lo->ny) / ((double)hi->nx * hi->ny);
  const size_t n = (size_t)lo->M * lo->nz * lo->nplane_c;
  hipLaunchKernelGGL(k_coarsen, dim3(grid_1d(n)), dim3(256), 0, lo->stream,
                     spectral_ptr(hi, field_hi), spectral_ptr(lo, field_lo), G, lo->M * lo->nz,
                     hi->nx, lo->nx, s);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(lo->stream));
  if (G) hipFree(G);
  if (field_lo == QGK_QH) {
    lo->primed = false;
    lo->q_stale = true;
  }
  return 0;
}

int qgk_subgrid_forcing(qgk_bucket* hi, qgk_bucket* lo, const double* filter_lo) {
  if (!hi || !lo) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(hi, "qgk_subgrid_forcing")) return rc0;
  if (!(hi->flags & QGK_FULL_STATE) || !(lo->flags & QGK_FULL_STATE) || !lo->field_r[QGK_DQ])
    return fail(QGK_E_STATE,
                "subgrid forcing needs full-state handles and dq on the coarse one");
  int rc;
  // high-resolution tendency (advection + beta; friction is linear and cancels)
  if ((rc = qgk_invert(hi)) || (rc = qgk_advect(hi))) return rc;
  // coarse state and its own tendency
  if ((rc = qgk_coarsen(hi, lo, QGK_QH, QGK_QH, filter_lo)) || (rc = refresh_q(lo)) ||
      (rc = qgk_invert(lo)) || (rc = qgk_advect(lo)))
    return rc;
  // filtered high-resolution tendency into dqh, then S_h = C(T_hi) - T_lo(C q)
  if ((rc = qgk_coarsen(hi, lo, QGK_DQHDT, QGK_DQH, filter_lo))) return rc;
  const size_t n = (size_t)lo->M * lo->nz * lo->nplane_c;
  hipLaunchKernelGGL(k_csub, dim3(grid_1d(n)), dim3(256), 0, lo->stream, lo->field_c[QGK_DQH],
                     D(lo), lo->field_c[QGK_DQH], n);
  HIPC(hipGetLastError());
  rc = c2r2d(lo, lo->field_c[QGK_DQH], lo->field_r[QGK_DQ], lo->M * lo->nz);
  if (rc) return rc;
  HIPC(hipStreamSynchronize(lo->stream));
  return 0;
}

// Upload n constants to dev only when they differ from the cached copy; under
// graph capture an upload would read a host buffer at replay, so it is refused.
int upload_cached(qgk_bucket* b, const double* host, int n, std::vector<double>& cache,
                  double* dev) {
  if (cache.size() == (size_t)n && std::equal(host, host + n, cache.begin())) return 0;
  if (b->capturing) return fail(QGK_E_STATE, "closure scales changed during graph capture");
  cache.assign(host, host + n);
  HIPC(hipMemcpyAsync(dev, cache.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
  return 0;
}

void* qgk_stream(qgk_bucket* b) { return b ? (void*)b->stream : nullptr; }

int qgk_pack_inputs(qgk_bucket* b, const double* inv_scale, int cinp, void* dst) {
  if (!b || !inv_scale || !dst) return fail(QGK_E_ARG, "NULL argument");
  if (!(b->flags & QGK_FULL_STATE)) return fail(QGK_E_STATE, "needs a full-state handle");
  if (cinp < 3 * b->nz) return fail(QGK_E_ARG, "cinp too small");
  int rc = refresh_q(b);
  if (rc) return rc;
  if (3 * b->nz > 128) return fail(QGK_E_ARG, "too many layers for the packing constants");
  double* sc = b->small;
  if ((rc = upload_cached(b, inv_scale, 3 * b->nz, b->small_in_h, sc))) return rc;
  if (cinp % 8) return fail(QGK_E_ARG, "cinp must be a multiple of 8");
  if (reinterpret_cast<uintptr_t>(dst) % 16) return fail(QGK_E_ARG, "dst must be 16-byte aligned");
  const int npix = (int)b->nplane_r;
  const size_t n = (size_t)b->M * npix * (cinp / 8);
  hipLaunchKernelGGL(k_pack_nhwc, dim3(grid_1d(n)), dim3(256), 0, b->stream, b->field_r[QGK_Q],
                     b->field_r[QGK_U], b->field_r[QGK_V], sc, (unsigned short*)dst, b->M, b->nz,
                     npix, cinp);
  HIPC(hipGetLastError());
  HIPC(stream_done(b));
  return 0;
}

// the stochastic forcing, when set, added to dqh after the transform of the
// network's forcing (both forcing paths)
static int apply_noise(qgk_bucket* b) {
  if (!b->noise_on) return 0;
  const double phi = b->noise_phi, sig = std::sqrt(1.0 - phi * phi);
  const size_t n = (size_t)b->M * b->nz * b->ny * b->nk;
  hipLaunchKernelGGL(k_spectral_noise, dim3(grid_1d(n)), dim3(256), 0, b->stream,
                     b->field_c[QGK_DQH], b->noise_xi, b->noise_amp, phi, sig, b->noise_seed,
                     b->noise_ctr, b->M, b->nz, b->ny, b->nx, b->nk, b->ldk);
  hipLaunchKernelGGL(k_noise_count, dim3(1), dim3(64), 0, b->stream, b->noise_ctr);
  HIPC(hipGetLastError());
  return 0;
}

int qgk_set_spectral_noise(qgk_bucket* b, const double* amp, double phi,
                           unsigned long long seed) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_set_spectral_noise")) return rc0;
  if (!amp) {
    b->noise_on = false;
    b->gen++;
    return 0;
  }
  if (!(phi >= 0.0 && phi < 1.0)) return fail(QGK_E_ARG, "phi must be in [0, 1)");
  if (!b->field_c[QGK_DQH])
    return fail(QGK_E_STATE, "dqh is not allocated (no q parameterization)");
  const size_t na = (size_t)b->nz * b->ny * b->nk;
  for (size_t i = 0; i < na; ++i)
    if (!(amp[i] >= 0.0) || !std::isfinite(amp[i]))
      return fail(QGK_E_ARG, "amplitudes must be finite and non-negative");
  int rc;
  if (!b->noise_xi) {
    if ((rc = dev_alloc(b, (void**)&b->noise_xi, (size_t)b->M * b->nz * b->nplane_c * sizeof(cd))))
      return rc;
    if ((rc = dev_alloc(b, (void**)&b->noise_amp, na * sizeof(double)))) return rc;
    if ((rc = dev_alloc(b, (void**)&b->noise_ctr, sizeof(unsigned long long)))) return rc;
  } else {  // a new process starts from zero
    HIPC(hipMemsetAsync(b->noise_xi, 0, (size_t)b->M * b->nz * b->nplane_c * sizeof(cd),
                        b->stream));
    HIPC(hipMemsetAsync(b->noise_ctr, 0, sizeof(unsigned long long), b->stream));
  }
  HIPC(hipMemcpyAsync(b->noise_amp, amp, na * sizeof(double), hipMemcpyHostToDevice, b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  b->noise_on = true;
  b->noise_phi = phi;
  b->noise_seed = seed;
  b->gen++;  // captured steps are recaptured
  return 0;
}

int qgk_add_forcing_device(qgk_bucket* b, const float* net_out, int coutp, const double* scale) {
  if (!b || !net_out || !scale) return fail(QGK_E_ARG, "NULL argument");
  if (!b->field_r[QGK_DQ]) return fail(QGK_E_STATE, "dq is not allocated (no q parameterization)");
  if (coutp < b->nz) return fail(QGK_E_ARG, "coutp too small");
  if (b->nz > 128) return fail(QGK_E_ARG, "too many layers for the forcing scales");
  double* sc = b->small + 128;
  int rc = upload_cached(b, scale, b->nz, b->small_out_h, sc);
  if (rc) return rc;
  const int np = b->M * b->nz;
  const size_t n = (size_t)np * b->nplane_r;
  hipLaunchKernelGGL(k_unpack_forcing, dim3(grid_1d(n)), dim3(256), 0, b->stream, net_out, sc,
                     b->field_r[QGK_DQ], b->M, b->nz, (int)b->nplane_r, coutp);
  HIPC(hipGetLastError());
  rc = r2c2d(b, b->field_r[QGK_DQ], b->field_c[QGK_DQH], np);
  if (rc) return rc;
  if ((rc = apply_noise(b))) return rc;
  const size_t cpts = (size_t)np * b->nplane_c;
  hipLaunchKernelGGL(k_add, dim3(grid_1d(cpts)), dim3(256), 0, b->stream, D(b),
                     b->field_c[QGK_DQH], cpts);
  HIPC(hipGetLastError());
  HIPC(stream_done(b));
  return 0;
}

int qgk_fused_row_net(qgk_bucket* b, const double* inv_scale, int cinp, void* net_in) {
  if (!b || !inv_scale || !net_in) return fail(QGK_E_ARG, "NULL argument");
  if (b->nz != 2 || b->nx != b->ny || b->nx > kFusedNetMaxN)
    return fail(QGK_E_ARG, "the fused closed step needs a two-layer square grid of n <= %d",
                kFusedNetMaxN);
  if (cinp < 3 * b->nz || cinp % 8) return fail(QGK_E_ARG, "cinp must be >= 6 and a multiple of 8");
  if (reinterpret_cast<uintptr_t>(net_in) % 16) return fail(QGK_E_ARG, "net_in must be 16-byte aligned");
  int rc = upload_params(b);
  if (rc) return rc;
  if (!packed_exchange(b))
    return fail(QGK_E_STATE, "the fused closed step needs the packed register kernels "
                             "(layer split off, standard wavenumber grid)");
  double* sc = b->small;
  if ((rc = upload_cached(b, inv_scale, 3 * b->nz, b->small_in_h, sc))) return rc;
  if (!b->primed) {
    if ((rc = launch_col_prologue(b))) return rc;
    b->primed = true;
  }
  rc = with_ne<RowFluxNetL>(b->nx, b->erow, b, (const double*)sc,
                            static_cast<unsigned short*>(net_in), cinp);
  if (rc) return rc;
  HIPC(hipGetLastError());
  if (b->capturing) b->ug_fused = true;
  HIPC(stream_done(b));
  return 0;
}

int qgk_fused_col_forcing(qgk_bucket* b, const float* net_out, int coutp, const double* scale,
                          double dt1, double dt2, double dt3) {
  if (!b || !net_out || !scale) return fail(QGK_E_ARG, "NULL argument");
  if (!b->field_r[QGK_DQ]) return fail(QGK_E_STATE, "dq is not allocated (no q parameterization)");
  if (b->nz != 2 || b->nx != b->ny || b->nx > kFusedNetMaxN)
    return fail(QGK_E_ARG, "the fused closed step needs a two-layer square grid of n <= %d",
                kFusedNetMaxN);
  if (coutp < b->nz) return fail(QGK_E_ARG, "coutp too small");
  if (!b->primed || !packed_exchange(b))
    return fail(QGK_E_STATE, "qgk_fused_row_net must begin the step");
  double* sc = b->small + 128;
  int rc = upload_cached(b, scale, b->nz, b->small_out_h, sc);
  if (rc) return rc;
  const int np = b->M * b->nz;
  const size_t n = (size_t)np * b->nplane_r;
  hipLaunchKernelGGL(k_unpack_forcing, dim3(grid_1d(n)), dim3(256), 0, b->stream, net_out, sc,
                     b->field_r[QGK_DQ], b->M, b->nz, (int)b->nplane_r, coutp);
  HIPC(hipGetLastError());
  if ((rc = r2c2d(b, b->field_r[QGK_DQ], b->field_c[QGK_DQH], np))) return rc;
  if ((rc = apply_noise(b))) return rc;
  const StepCoeffs cf = {dt1, dt2, dt3};
  if ((rc = with_ne<ColStepForceL>(b->ny, b->ecol, b, cf))) return rc;
  HIPC(hipGetLastError());
  rotate_history(b);
  std::swap(b->field_c[QGK_QH], b->qh_alt);
  b->q_stale = true;
  if (b->capturing) b->ug_fused = true;
  HIPC(stream_done(b));
  return 0;
}

int qgk_step(qgk_bucket* b, int nsteps, const double* coeffs, int use_graph) {
  return qgk_step_buckets(&b, 1, nsteps, coeffs, use_graph);
}

int qgk_member_speed(qgk_bucket* b, double* speed_host) {
  if (!b || !speed_host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_member_speed")) return rc0;
  int rc = 0;
  if (!b->speed && (rc = dev_alloc(b, (void**)&b->speed, b->M * sizeof(unsigned long long))))
    return rc;
  if ((rc = launch_member_speed(b, b->speed))) return rc;
  std::vector<unsigned long long> w(b->M);
  HIPC(hipMemcpyAsync(w.data(), b->speed, b->M * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost, b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  for (int m = 0; m < b->M; ++m) std::memcpy(speed_host + m, &w[m], sizeof(double));
  return 0;
}

int qgk_member_speed_noalloc(qgk_bucket* b, double* speed_dev) {
  if (!b || !speed_dev) return fail(QGK_E_ARG, "NULL argument");
  if (b->M > kNoAllocMaxMembers)
    return fail(QGK_E_ARG, "qgk_member_speed_noalloc takes at most %d members (bucket has %d)",
                kNoAllocMaxMembers, b->M);
  unsigned long long* w = nullptr;
  int rc = scratch_address(&w);
  if (rc || (rc = launch_member_speed(b, w))) return rc;
  hipLaunchKernelGGL(k_u64_to_double, dim3((b->M + 255) / 256), dim3(256), 0, b->stream, w,
                     speed_dev, b->M);
  HIPC(hipGetLastError());
  HIPC(stream_done(b));
  return 0;
}

int qgk_mask_member(qgk_bucket* b, int member) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "qgk_mask_member")) return rc0;
  if (member < 0 || member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  const size_t cm = (size_t)b->nz * b->nplane_c, rm = (size_t)b->nz * b->nplane_r;
  cd* cs[] = {b->field_c[QGK_QH], b->qh_alt, b->hist[0], b->hist[1], b->hist[2], b->Qm, b->Um,
              b->Vm, b->UQm, b->VQm, b->phm, b->field_c[QGK_PH], b->field_c[QGK_UH],
              b->field_c[QGK_VH], b->field_c[QGK_UQH], b->field_c[QGK_VQH],
              b->field_c[QGK_DQH], b->field_c[QGK_DUH], b->field_c[QGK_DVH]};
  double* rs[] = {b->field_r[QGK_Q], b->field_r[QGK_U], b->field_r[QGK_V], b->field_r[QGK_UQ],
                  b->field_r[QGK_VQ], b->field_r[QGK_DQ], b->field_r[QGK_DU],
                  b->field_r[QGK_DV]};
  for (cd* f : cs)
    if (f) HIPC(hipMemsetAsync(f + member * cm, 0, cm * sizeof(cd), b->stream));
  for (double* f : rs)
    if (f) HIPC(hipMemsetAsync(f + member * rm, 0, rm * sizeof(double), b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"

// ==========================================================================
// Diagnostics (host side)
// ==========================================================================
namespace {

bool diag_is_scalar(int d) { return d == QGK_D_EKEDISS || d == QGK_D_EKE || d == QGK_D_APEGEN; }

// Slot table for this bucket's nz; -1 marks diagnostics that do not exist for it.
void diag_layout(qgk_bucket* b) {
  if (b->nslot || b->nscal) return;
  const int nz = b->nz;
  for (int d = 0; d < QGK_NDIAG; ++d) {
    int n = 1;
    if (d == QGK_D_ENSSPEC || d == QGK_D_KESPEC || d == QGK_D_KESPEC_MODAL || d == QGK_D_EKE)
      n = nz;
    if (d == QGK_D_PESPEC_MODAL) n = nz - 1;
    bool ok = true;
    if ((d == QGK_D_APEFLUX || d == QGK_D_KEFLUX || d == QGK_D_APEGEN) && nz != 2) ok = false;
    if ((d == QGK_D_PESPEC_MODAL || d == QGK_D_APESPEC) && nz < 2) ok = false;
    if (!ok || n <= 0) {
      b->dslot[d] = -1;
      b->dnsl[d] = 0;
      continue;
    }
    b->dnsl[d] = n;
    if (diag_is_scalar(d)) {
      b->dslot[d] = b->nscal;
      b->nscal += n;
    } else {
      b->dslot[d] = b->nslot;
      b->nslot += n;
    }
  }
}

int diag_plen(const qgk_bucket* b) { return DiagLayout{b->nz}.len(); }

// Work space shared by the diagnostics and the built-in closures: 5 spectral
// and 4 real sets of M*nz planes.
int ensure_work(qgk_bucket* b) {
  if (b->dW) return 0;
  const size_t np = (size_t)b->M * b->nz;
  int rc;
  if ((rc = dev_alloc(b, (void**)&b->dW, 5 * np * b->nplane_c * sizeof(cd))) ||
      (rc = dev_alloc(b, (void**)&b->dR, 4 * np * b->nplane_r * sizeof(double))))
    return rc;
  return 0;
}

bool diag_needs_jacobian(unsigned mask) {
  const unsigned j = (1u << QGK_D_ENSFLUX) | (1u << QGK_D_KEFLUX) | (1u << QGK_D_APEFLUX) |
                     (1u << QGK_D_KEFLUX_DIV) | (1u << QGK_D_APEFLUX_DIV);
  return (mask & j) != 0;
}

// Jacobian sign * (ik fft(u x) + il fft(v x)) of the real planes x (nzx per
// member) into J; mode 1 combines the two layers' velocities.
int diag_advect(qgk_bucket* b, const double* x, int nzx, int mode, double sign, cd* J) {
  const size_t np = (size_t)b->M * b->nz;
  cd* W1 = b->dW;
  cd* W2 = b->dW + np * b->nplane_c;
  double* P1 = b->dR + np * b->nplane_r;
  double* P2 = b->dR + 2 * np * b->nplane_r;
  const int nplanes = b->M * nzx;
  const size_t nr = (size_t)nplanes * b->nplane_r;
  hipLaunchKernelGGL(k_diag_products, dim3(grid_1d(nr)), dim3(256), 0, b->stream, x,
                     b->field_r[QGK_U], b->field_r[QGK_V], b->dprm, P1, P2, b->M, b->nz,
                     (int)b->nplane_r, mode, diag_plen(b));
  HIPC(hipGetLastError());
  int rc;
  {
    const double* in[2] = {P1, P2};
    cd* out[2] = {W1, W2};
    if ((rc = r2c2d_n(b, 2, in, out, nplanes))) return rc;
  }
  const size_t nc = (size_t)nplanes * b->nplane_c;
  hipLaunchKernelGGL(k_diag_jacobian, dim3(grid_1d(nc)), dim3(256), 0, b->stream, W1, W2,
                     b->dp, J, b->M, nzx, b->nl, b->ldk, sign);
  HIPC(hipGetLastError());
  return 0;
}

// inverse transform of a prepared spectral input (k_diag_prepare mode) into x
int diag_prepare_real(qgk_bucket* b, int mode, double* x) {
  const int nzx = mode == 1 ? 1 : b->nz;
  const size_t nc = (size_t)b->M * nzx * b->nplane_c;
  hipLaunchKernelGGL(k_diag_prepare, dim3(grid_1d(nc)), dim3(256), 0, b->stream,
                     b->field_c[QGK_PH], b->dp, b->dprm, b->dW, b->M, b->nz, b->nl, b->ldk, mode,
                     diag_plen(b));
  HIPC(hipGetLastError());
  return c2r2d(b, b->dW, x, b->M * nzx);
}

hipError_t copy_real_rows(qgk_bucket* b, double* dst, const double* src, size_t rows) {
  return hipMemcpy2DAsync(dst, (size_t)b->nk * sizeof(double), src,
                          (size_t)b->ldk * sizeof(double), (size_t)b->nk * sizeof(double), rows,
                          hipMemcpyDeviceToHost, b->stream);
}

}  // namespace

extern "C" {

int qgk_diag_set_params(qgk_bucket* b, int member, const double* params) {
  if (!b || !params) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "diagnostics")) return rc0;
  if (member < 0 || member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  if (b->nz > kDiagNzMax) return fail(QGK_E_ARG, "diagnostics support nz <= %d", kDiagNzMax);
  const int plen = diag_plen(b);
  if (b->dprm_h.empty()) {
    b->dprm_h.assign((size_t)b->M * plen, 0.0);
    b->dprm_set.assign(b->M, 0);
  }
  double* dst = b->dprm_h.data() + (size_t)member * plen;
  if (!b->dprm_set[member] || std::memcmp(dst, params, plen * sizeof(double)) != 0) {
    // an accumulation may still be copying the table (accumulations do not wait)
    HIPC(hipStreamSynchronize(b->stream));
    std::memcpy(dst, params, plen * sizeof(double));
    b->dprm_dirty = true;  // uploaded with the next accumulation
  }
  b->dprm_set[member] = 1;
  return 0;
}

int qgk_diag_reset(qgk_bucket* b) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "diagnostics")) return rc0;
  const size_t plane = (size_t)b->nl * b->ldk;
  if (b->dacc)
    HIPC(hipMemsetAsync(b->dacc, 0, (size_t)b->nslot * b->M * plane * sizeof(double), b->stream));
  if (b->dsc) HIPC(hipMemsetAsync(b->dsc, 0, (size_t)b->nscal * b->M * sizeof(double), b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_diag_accumulate(qgk_bucket* b, unsigned int mask, int instant, double dt, double dt1,
                        double dt2, double dt3) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  if (int rc0 = refuse_in_capture(b, "diagnostics")) return rc0;
  int rc = need_full(b, "diagnostics");
  if (rc) return rc;
  if (b->nz > kDiagNzMax) return fail(QGK_E_ARG, "diagnostics support nz <= %d", kDiagNzMax);
  diag_layout(b);
  if (mask >> QGK_NDIAG) return fail(QGK_E_ARG, "unknown diagnostic bits in mask 0x%x", mask);
  for (int d = 0; d < QGK_NDIAG; ++d)
    if ((mask >> d & 1u) && b->dslot[d] < 0)
      return fail(QGK_E_ARG, "diagnostic %d does not exist for nz=%d", d, b->nz);
  if ((mask >> QGK_D_APEFLUX & 1u) && (mask >> QGK_D_APEFLUX_DIV & 1u))
    return fail(QGK_E_ARG, "APEflux and APEflux_div cannot be accumulated in one call");
  if (dt == 0.0) return fail(QGK_E_ARG, "dt must be nonzero");
  if (b->dprm_set.empty()) return fail(QGK_E_STATE, "diagnostic parameters not set");
  for (int m = 0; m < b->M; ++m)
    if (!b->dprm_set[m])
      return fail(QGK_E_STATE, "diagnostic parameters not set for member %d", m);
  if (mask == 0) return 0;
  const size_t plane = (size_t)b->nl * b->ldk;
  const size_t np = (size_t)b->M * b->nz;
  const int plen = diag_plen(b);
  // buffers
  if (!b->dprm && (rc = dev_alloc(b, (void**)&b->dprm, b->dprm_h.size() * sizeof(double))))
    return rc;
  if (!b->dacc) {
    if ((rc = dev_alloc(b, (void**)&b->dacc, (size_t)b->nslot * b->M * plane * sizeof(double))) ||
        (rc = dev_alloc(b, (void**)&b->dsc, (size_t)(b->nscal + 1) * b->M * sizeof(double))) ||
        (rc = dev_alloc(b, (void**)&b->dsum,
                        ((size_t)b->M * (b->nz + 1) + (size_t)b->nz * plane) * sizeof(double))))
      return rc;
  }
  if (instant && !b->dacc_i) {
    if ((rc = dev_alloc(b, (void**)&b->dacc_i, (size_t)b->nslot * b->M * plane * sizeof(double))) ||
        (rc = dev_alloc(b, (void**)&b->dsc_i, (size_t)(b->nscal + 1) * b->M * sizeof(double))))
      return rc;
  }
  const bool apegen = mask >> QGK_D_APEGEN & 1u;
  if ((diag_needs_jacobian(mask) || apegen) && (rc = ensure_work(b))) return rc;
  if (b->dprm_dirty) {
    HIPC(hipMemcpyAsync(b->dprm, b->dprm_h.data(), b->dprm_h.size() * sizeof(double),
                        hipMemcpyHostToDevice, b->stream));
    b->dprm_dirty = false;
  }
  if ((rc = upload_params(b)) || (rc = refresh_q(b))) return rc;
  double* acc = instant ? b->dacc_i : b->dacc;
  double* accs = instant ? b->dsc_i : b->dsc;
  if (instant) {
    HIPC(hipMemsetAsync(acc, 0, (size_t)b->nslot * b->M * plane * sizeof(double), b->stream));
    HIPC(hipMemsetAsync(accs, 0, (size_t)b->nscal * b->M * sizeof(double), b->stream));
  }
  // Jacobians of the derived fields (qg_model.py:213-224, layered_model.py:242-254)
  cd* Jq = b->dW + 2 * np * b->nplane_c;
  cd* Jpxi = b->dW + 3 * np * b->nplane_c;
  cd* J3 = b->dW + 4 * np * b->nplane_c;
  double* X = b->dR;
  if (mask >> QGK_D_ENSFLUX & 1u)
    if ((rc = diag_advect(b, b->field_r[QGK_Q], b->nz, 0, 1.0, Jq))) return rc;
  if ((mask >> QGK_D_KEFLUX & 1u) || (mask >> QGK_D_KEFLUX_DIV & 1u))
    if ((rc = diag_prepare_real(b, 0, X)) || (rc = diag_advect(b, X, b->nz, 0, 1.0, Jpxi)))
      return rc;
  if (mask >> QGK_D_APEFLUX & 1u)
    if ((rc = diag_prepare_real(b, 1, X)) || (rc = diag_advect(b, X, 1, 1, -1.0, J3))) return rc;
  if (mask >> QGK_D_APEFLUX_DIV & 1u)
    if ((rc = diag_prepare_real(b, 2, X)) || (rc = diag_advect(b, X, b->nz, 0, 1.0, J3)))
      return rc;
  DiagArgs A;
  A.qh = b->field_c[QGK_QH];
  A.ph = b->field_c[QGK_PH];
  A.D = D(b);
  A.P = Pp(b);
  A.PP = PP(b);
  A.dqh = (b->flags & QGK_HAS_Q_PARAM) ? b->field_c[QGK_DQH] : nullptr;
  A.duh = (b->flags & QGK_HAS_UV_PARAM) ? b->field_c[QGK_DUH] : nullptr;
  A.dvh = (b->flags & QGK_HAS_UV_PARAM) ? b->field_c[QGK_DVH] : nullptr;
  A.Jq = Jq;
  A.Jpxi = Jpxi;
  A.J3 = J3;
  A.prm = b->dprm;
  A.acc = acc;
  A.pts = apegen ? b->dR + np * b->nplane_r : nullptr;
  A.M = b->M;
  A.nz = b->nz;
  A.nl = b->nl;
  A.nk = b->nk;
  A.ldk = b->ldk;
  A.plen = plen;
  const double M2 = (double)b->nx * (double)b->ny;
  A.invM2 = 1.0 / (M2 * M2);
  A.dt = dt;
  A.dt1 = dt1;
  A.dt2 = dt2;
  A.dt3 = dt3;
  for (int d = 0; d < QGK_NDIAG; ++d)
    A.slot[d] = (!diag_is_scalar(d) && (mask >> d & 1u)) ? b->dslot[d] : -1;
  const size_t total = (size_t)b->M * plane;
  if (b->nz <= kDiagNzSmall)
    hipLaunchKernelGGL(k_diag_spectral<kDiagNzSmall>, dim3(grid_1d(total)), dim3(256), 0,
                       b->stream, A, b->dp);
  else
    hipLaunchKernelGGL(k_diag_spectral<kDiagNzMax>, dim3(grid_1d(total)), dim3(256), 0,
                       b->stream, A, b->dp);
  HIPC(hipGetLastError());
  const bool eke = (mask >> QGK_D_EKE & 1u) || (mask >> QGK_D_EKEDISS & 1u);
  if (eke) {
    hipLaunchKernelGGL(k_diag_plane_sum, dim3((unsigned)np), dim3(256), 0, b->stream,
                       b->field_r[QGK_U], b->field_r[QGK_V], b->dsum, b->nplane_r, 1);
    HIPC(hipGetLastError());
  }
  if (apegen) {
    hipLaunchKernelGGL(k_diag_plane_sum, dim3((unsigned)b->M), dim3(256), 0, b->stream, A.pts,
                       (const double*)nullptr, b->dsum + np, plane, 0);
    HIPC(hipGetLastError());
  }
  if (eke || apegen) {
    hipLaunchKernelGGL(k_diag_scalars, dim3((b->M + 63) / 64), dim3(64), 0, b->stream, b->dsum,
                       b->dprm, b->dp, accs, b->M, b->nz, plen, (double)b->nplane_r, A.invM2,
                       (mask >> QGK_D_EKE & 1u) ? b->dslot[QGK_D_EKE] : -1,
                       (mask >> QGK_D_EKEDISS & 1u) ? b->dslot[QGK_D_EKEDISS] : -1,
                       apegen ? b->dslot[QGK_D_APEGEN] : -1);
    HIPC(hipGetLastError());
  }
  // no wait: the sums stay on the device until qgk_diag_read, which waits
  HIPC(hipGetLastError());
  return 0;
}

int qgk_diag_read(qgk_bucket* b, int diag, int member, int instant, double* host) {
  if (!b || !host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "diagnostics")) return rc0;
  if (diag < 0 || diag >= QGK_NDIAG) return fail(QGK_E_ARG, "bad diagnostic id %d", diag);
  diag_layout(b);
  if (b->dslot[diag] < 0)
    return fail(QGK_E_ARG, "diagnostic %d does not exist for nz=%d", diag, b->nz);
  if (member < -2 || member >= b->M) return fail(QGK_E_ARG, "member %d out of range", member);
  double* acc = instant ? b->dacc_i : b->dacc;
  double* accs = instant ? b->dsc_i : b->dsc;
  if (!acc) return fail(QGK_E_STATE, "no diagnostics accumulated");
  const int nsl = b->dnsl[diag], s0 = b->dslot[diag];
  const size_t plane = (size_t)b->nl * b->ldk;
  if (diag_is_scalar(diag)) {
    std::vector<double> all((size_t)nsl * b->M);
    HIPC(hipMemcpyAsync(all.data(), accs + (size_t)s0 * b->M, all.size() * sizeof(double),
                        hipMemcpyDeviceToHost, b->stream));
    HIPC(hipStreamSynchronize(b->stream));
    for (int j = 0; j < nsl; ++j) {
      if (member >= 0) {
        host[j] = all[(size_t)j * b->M + member];
      } else if (member == -1) {
        for (int m = 0; m < b->M; ++m) host[(size_t)m * nsl + j] = all[(size_t)j * b->M + m];
      } else {
        double s = 0.0;
        for (int m = 0; m < b->M; ++m) s += all[(size_t)j * b->M + m];
        host[j] = s;
      }
    }
    return 0;
  }
  const size_t dense = (size_t)b->nl * b->nk;
  if (member >= 0) {
    for (int j = 0; j < nsl; ++j)
      HIPC(copy_real_rows(b, host + j * dense, acc + ((size_t)(s0 + j) * b->M + member) * plane,
                          b->nl));
  } else if (member == -1) {
    for (int m = 0; m < b->M; ++m)
      for (int j = 0; j < nsl; ++j)
        HIPC(copy_real_rows(b, host + ((size_t)m * nsl + j) * dense,
                            acc + ((size_t)(s0 + j) * b->M + m) * plane, b->nl));
  } else {
    double* tmp = b->dsum + (size_t)b->M * (b->nz + 1);
    const size_t n = (size_t)nsl * plane;
    hipLaunchKernelGGL(k_diag_member_sum, dim3(grid_1d(n)), dim3(256), 0, b->stream,
                       acc + (size_t)s0 * b->M * plane, tmp, b->M, nsl, plane);
    HIPC(hipGetLastError());
    HIPC(copy_real_rows(b, host, tmp, (size_t)nsl * b->nl));
  }
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

}  // extern "C"

// ==========================================================================
// Built-in closures on the device (host side)
// ==========================================================================
namespace {

const size_t kGrid = 256;

int param_setup(qgk_bucket* b, const double* dx, const double* Hi, bool need_uv) {
  int rc = need_full(b, "built-in closures");
  if (rc) return rc;
  if (need_uv && !b->field_r[QGK_U]) return fail(QGK_E_STATE, "u/v not allocated");
  if ((rc = ensure_work(b)) || (rc = upload_params(b)) || (rc = refresh_q(b))) return rc;
  const size_t np = (size_t)b->M * b->nz;
  if (!b->pdx && ((rc = dev_alloc(b, (void**)&b->pdx, b->M * sizeof(double))) ||
                  (rc = dev_alloc(b, (void**)&b->pHi, np * sizeof(double))) ||
                  (rc = dev_alloc(b, (void**)&b->psum, 2 * np * sizeof(double)))))
    return rc;
  // upload dx and Hi only when they change; under graph capture an upload
  // would become a copy from a host buffer that is gone at replay
  auto upload = [&](const double* host, std::vector<double>& cache, double* dev,
                    size_t n) -> int {
    if (!host || (cache.size() == n && std::equal(host, host + n, cache.begin()))) return 0;
    if (b->capturing) return fail(QGK_E_STATE, "closure constants changed during graph capture");
    cache.assign(host, host + n);
    HIPC(hipMemcpyAsync(dev, cache.data(), n * sizeof(double), hipMemcpyHostToDevice, b->stream));
    return 0;
  };
  if ((rc = upload(dx, b->pdx_h, b->pdx, b->M)) || (rc = upload(Hi, b->pHi_h, b->pHi, np)))
    return rc;
  return 0;
}

cd* Wset(qgk_bucket* b, int i) { return b->dW + (size_t)i * b->M * b->nz * b->nplane_c; }
double* Rset(qgk_bucket* b, int i) { return b->dR + (size_t)i * b->M * b->nz * b->nplane_r; }

int spec_ikil(qgk_bucket* b, const cd* X, const cd* Y, cd* out, double cx, double cy) {
  const size_t n = (size_t)b->M * b->nz * b->nplane_c;
  hipLaunchKernelGGL(k_spec_ikil, dim3(grid_1d(n)), dim3(kGrid), 0, b->stream, X, Y, out, b->dp,
                     b->M, b->nz, b->nl, b->ldk, cx, cy);
  HIPC(hipGetLastError());
  return 0;
}

int spec_wv2(qgk_bucket* b, const cd* X, cd* out, int power, double sign) {
  const size_t n = (size_t)b->M * b->nz * b->nplane_c;
  hipLaunchKernelGGL(k_spec_wv2, dim3(grid_1d(n)), dim3(kGrid), 0, b->stream, X, out, b->dp,
                     b->M, b->nz, b->nl, b->ldk, power, sign);
  HIPC(hipGetLastError());
  return 0;
}

// Inverse transforms of nf spectral operator results (SpecJobs kinds; M*nz
// planes each) into out[f]: formed inside the one-launch plane transform where
// that applies, else by spec_ikil / spec_wv2 into work[f] and the row and
// column kernels.
int c2r2d_spec(qgk_bucket* b, int nf, const SpecJobs& ops, double* const* out,
               cd* const* work) {
  if (nf < 1 || nf > kPlaneJobs) return fail(QGK_E_ARG, "1 to %d fields per transform", kPlaneJobs);
  const int np = b->M * b->nz;
  const size_t ib = (size_t)np * b->nplane_c * sizeof(cd);
  const size_t ob = (size_t)np * b->nplane_r * sizeof(double);
  bool plane = plane_ok(b);
  for (int f = 0; plane && f < nf; ++f)
    for (int g = 0; g < nf; ++g) {
      const void* srcs[2] = {ops.X[g], ops.Y[g]};
      for (const void* src : srcs)
        if (src && jobs_overlap(1, &src, ib, reinterpret_cast<void* const*>(&out[f]), ob))
          plane = false;
    }
  if (plane) {
    PlaneJobs j{};
    for (int f = 0; f < nf; ++f) j.out[f] = out[f];
    return launch_plane<PlaneSpecL>(b, j, nf, np, ops);
  }
  const cd* in[kPlaneJobs];
  for (int f = 0; f < nf; ++f) {
    int rc = 0;
    if (ops.kind[f] == kSpecPlain) {
      in[f] = ops.X[f];
      continue;
    }
    if (ops.kind[f] == kSpecIkil)
      rc = spec_ikil(b, ops.X[f], ops.Y[f], work[f], ops.cx[f], ops.cy[f]);
    else
      rc = spec_wv2(b, ops.X[f], work[f], ops.power[f], ops.cx[f]);
    if (rc) return rc;
    in[f] = work[f];
  }
  return c2r2d_n(b, nf, in, out, np);
}

// SpecJobs of nf ik/il derivative fields
SpecJobs ikil_jobs(int nf, const cd* const* X, const cd* const* Y, const double* cx,
                   const double* cy) {
  SpecJobs o{};
  for (int f = 0; f < nf; ++f) {
    o.X[f] = X[f];
    o.Y[f] = Y[f];
    o.cx[f] = cx[f];
    o.cy[f] = cy[f];
    o.kind[f] = kSpecIkil;
  }
  return o;
}

int c2r2d_ikil(qgk_bucket* b, int nf, const cd* const* X, const cd* const* Y, const double* cx,
               const double* cy, double* const* out, cd* const* work) {
  return c2r2d_spec(b, nf, ikil_jobs(nf, X, Y, cx, cy), out, work);
}

// weight * scale * src into the field dst (or added to it), or to host memory
int param_emit(qgk_bucket* b, double* src, double* dst, double* host, double scale,
               double weight, int accumulate) {
  const size_t n = (size_t)b->M * b->nz * b->nplane_r;
  if (host) {
    hipLaunchKernelGGL(k_param_out, dim3(grid_1d(n)), dim3(kGrid), 0, b->stream, src, src, scale,
                       weight, 0, n);
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(host, src, n * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    return 0;
  }
  hipLaunchKernelGGL(k_param_out, dim3(grid_1d(n)), dim3(kGrid), 0, b->stream, src, dst, scale,
                     weight, accumulate, n);
  HIPC(hipGetLastError());
  return 0;
}

// Smagorinsky first stage (parameterizations.py:174-181): R3 = nu and
// R0, R1, R2 = nu Sxx, nu Syy, nu Sxy
int smag_viscosity(qgk_bucket* b, double cs) {
  const int np = b->M * b->nz;
  cd *W0 = Wset(b, 0), *W1 = Wset(b, 1), *W2 = Wset(b, 2), *W3 = Wset(b, 3), *W4 = Wset(b, 4);
  double *R0 = Rset(b, 0), *R1 = Rset(b, 1), *R2 = Rset(b, 2), *R3 = Rset(b, 3);
  const double* uv[2] = {b->field_r[QGK_U], b->field_r[QGK_V]};
  cd* uvh[2] = {W0, W1};
  // ux = ifft(ik uh), vy = ifft(il vh), ifft(ik vh + il uh)
  const cd* X[3] = {W0, nullptr, W1};
  const cd* Y[3] = {nullptr, W1, W0};
  const double cx[3] = {1.0, 0.0, 1.0}, cy[3] = {0.0, 1.0, 1.0};
  double* Sr[3] = {R0, R1, R2};
  cd* work[3] = {W2, W3, W4};
  int rc;
  if ((rc = r2c2d_n(b, 2, uv, uvh, np)) || (rc = c2r2d_ikil(b, 3, X, Y, cx, cy, Sr, work)))
    return rc;
  const size_t n = (size_t)np * b->nplane_r;
  hipLaunchKernelGGL(k_smag_nu, dim3(grid_1d(n)), dim3(kGrid), 0, b->stream, R0, R1, R2, R3,
                     b->pdx, cs, b->M, b->nz, (int)b->nplane_r);
  HIPC(hipGetLastError());
  return 0;
}

int check_uv_target(qgk_bucket* b, const double* du_host, const double* dv_host) {
  if ((du_host == nullptr) != (dv_host == nullptr))
    return fail(QGK_E_ARG, "give both du_host and dv_host or neither");
  if (!du_host && !b->field_r[QGK_DU])
    return fail(QGK_E_STATE, "du/dv are not allocated (no uv parameterization)");
  return 0;
}

}  // namespace

extern "C" {

int qgk_param_smagorinsky(qgk_bucket* b, const double* dx, double constant, double weight,
                          int accumulate, double* du_host, double* dv_host) {
  if (!b || !dx) return fail(QGK_E_ARG, "NULL argument");
  int rc;
  if ((rc = check_uv_target(b, du_host, dv_host)) || (rc = param_setup(b, dx, nullptr, true)) ||
      (rc = smag_viscosity(b, constant)))
    return rc;
  const int np = b->M * b->nz;
  cd *W0 = Wset(b, 0), *W1 = Wset(b, 1), *W2 = Wset(b, 2), *W3 = Wset(b, 3), *W4 = Wset(b, 4);
  double *R0 = Rset(b, 0), *R1 = Rset(b, 1), *R2 = Rset(b, 2);
  const double* st[3] = {R0, R2, R1};
  cd* sth[3] = {W0, W1, W2};
  const cd* X[2] = {W0, W1};
  const cd* Y[2] = {W1, W2};
  const double one[2] = {1.0, 1.0};
  double* duv[2] = {R0, R1};
  cd* work[2] = {W3, W4};
  // du = 2 ifft(ik F(nu Sxx) + il F(nu Sxy)), dv = 2 ifft(ik F(nu Sxy) + il F(nu Syy))
  if ((rc = r2c2d_n(b, 3, st, sth, np)) || (rc = c2r2d_ikil(b, 2, X, Y, one, one, duv, work)) ||
      (rc = param_emit(b, R0, b->field_r[QGK_DU], du_host, 2.0, weight, accumulate)) ||
      (rc = param_emit(b, R1, b->field_r[QGK_DV], dv_host, 2.0, weight, accumulate)))
    return rc;
  HIPC(stream_done(b));
  return 0;
}

int qgk_smagorinsky_viscosity(qgk_bucket* b, const double* dx, double constant,
                              double* nu_host) {
  if (!b || !dx || !nu_host) return fail(QGK_E_ARG, "NULL argument");
  if (int rc0 = refuse_in_capture(b, "qgk_smagorinsky_viscosity")) return rc0;
  int rc;
  if ((rc = param_setup(b, dx, nullptr, true)) || (rc = smag_viscosity(b, constant))) return rc;
  const size_t n = (size_t)b->M * b->nz * b->nplane_r;
  HIPC(hipMemcpyAsync(nu_host, Rset(b, 3), n * sizeof(double), hipMemcpyDeviceToHost,
                      b->stream));
  HIPC(hipStreamSynchronize(b->stream));
  return 0;
}

int qgk_param_zb2020(qgk_bucket* b, double constant, double weight, int accumulate,
                     double* du_host, double* dv_host) {
  if (!b) return fail(QGK_E_ARG, "NULL handle");
  int rc;
  if ((rc = check_uv_target(b, du_host, dv_host)) ||
      (rc = param_setup(b, nullptr, nullptr, true)))
    return rc;
  const int np = b->M * b->nz;
  cd *W0 = Wset(b, 0), *W1 = Wset(b, 1), *W2 = Wset(b, 2), *W3 = Wset(b, 3), *W4 = Wset(b, 4);
  double *R0 = Rset(b, 0), *R1 = Rset(b, 1), *R2 = Rset(b, 2), *R3 = Rset(b, 3);
  const double* uv[2] = {b->field_r[QGK_U], b->field_r[QGK_V]};
  cd* uvh[2] = {W0, W1};
  // vx, vy, ux, uy (parameterizations.py:264-269)
  const cd* X[4] = {W1, nullptr, W0, nullptr};
  const cd* Y[4] = {nullptr, W1, nullptr, W0};
  const double cx[4] = {1.0, 0.0, 1.0, 0.0}, cy[4] = {0.0, 1.0, 0.0, 1.0};
  double* g4[4] = {R0, R1, R2, R3};
  cd* work[3] = {W2, W3, W4};
  if ((rc = r2c2d_n(b, 2, uv, uvh, np)) || (rc = c2r2d_ikil(b, 3, X, Y, cx, cy, g4, work)) ||
      (rc = c2r2d_ikil(b, 1, X + 3, Y + 3, cx + 3, cy + 3, g4 + 3, work)))
    return rc;
  const size_t nr = (size_t)np * b->nplane_r, nc = (size_t)np * b->nplane_c;
  hipLaunchKernelGGL(k_zb_products, dim3(grid_1d(nr)), dim3(kGrid), 0, b->stream, R0, R1, R2, R3,
                     nr);
  HIPC(hipGetLastError());
  // W0 = F(vort stretch), W1 = F(vort shear), W2 = F(sum of squares)
  {
    const double* pr[3] = {R0, R1, R2};
    cd* ph3[3] = {W0, W1, W2};
    if ((rc = r2c2d_n(b, 3, pr, ph3, np))) return rc;
  }
  // du = k ifft(ik (S/2 - shear) + il stretch); dv = k ifft(il (S/2 + shear) + ik stretch)
  hipLaunchKernelGGL(k_caxpby, dim3(grid_1d(nc)), dim3(kGrid), 0, b->stream, W2, W1, W3, 0.5,
                     -1.0, nc);
  HIPC(hipGetLastError());
  hipLaunchKernelGGL(k_caxpby, dim3(grid_1d(nc)), dim3(kGrid), 0, b->stream, W2, W1, W4, 0.5, 1.0,
                     nc);
  HIPC(hipGetLastError());
  const cd* dX[2] = {W3, W0};
  const cd* dY[2] = {W0, W4};
  const double one[2] = {1.0, 1.0};
  double* duv[2] = {R0, R1};
  cd* dwork[2] = {W3, W4};
  if ((rc = c2r2d_ikil(b, 2, dX, dY, one, one, duv, dwork)) ||
      (rc = param_emit(b, R0, b->field_r[QGK_DU], du_host, constant, weight, accumulate)) ||
      (rc = param_emit(b, R1, b->field_r[QGK_DV], dv_host, constant, weight, accumulate)))
    return rc;
  HIPC(stream_done(b));
  return 0;
}

int qgk_param_backscatter(qgk_bucket* b, const double* dx, const double* Hi,
                          double smag_constant, double back_constant, double eps, double weight,
                          int accumulate, double* dq_host) {
  if (!b || !dx || !Hi) return fail(QGK_E_ARG, "NULL argument");
  if (!dq_host && !b->field_r[QGK_DQ])
    return fail(QGK_E_STATE, "dq is not allocated (no q parameterization)");
  int rc;
  if ((rc = param_setup(b, dx, Hi, true)) || (rc = smag_viscosity(b, smag_constant))) return rc;
  const int np = b->M * b->nz;
  cd* W0 = Wset(b, 0);
  double *R0 = Rset(b, 0), *R1 = Rset(b, 1), *R2 = Rset(b, 2), *R3 = Rset(b, 3);
  const size_t nr = (size_t)np * b->nplane_r;
  // psi = ifft(ph), lap_lap_psi = ifft(lap^2 ph) (parameterizations.py:226-228)
  SpecJobs psi{};
  psi.X[0] = psi.X[1] = b->field_c[QGK_PH];
  psi.kind[0] = kSpecPlain;
  psi.kind[1] = kSpecWv2;
  psi.power[1] = 2;
  psi.cx[1] = 1.0;
  double* psir[2] = {R0, R1};
  cd* psiw[2] = {nullptr, W0};
  if ((rc = c2r2d_spec(b, 2, psi, psir, psiw))) return rc;
  // dissipation = -ifft(lap fft(lap_lap_psi dx^2 nu))
  hipLaunchKernelGGL(k_bs_product, dim3(grid_1d(nr)), dim3(kGrid), 0, b->stream, R1, R3, b->pdx,
                     R2, b->M, b->nz, (int)b->nplane_r);
  HIPC(hipGetLastError());
  SpecJobs lap{};
  lap.X[0] = W0;
  lap.kind[0] = kSpecWv2;
  lap.power[0] = 1;
  lap.cx[0] = -1.0;
  if ((rc = r2c2d(b, R2, W0, np)) || (rc = c2r2d_spec(b, 1, lap, &R2, &W0))) return rc;
  hipLaunchKernelGGL(k_plane_dot, dim3((unsigned)np), dim3(256), 0, b->stream, R0, R2, b->psum,
                     b->nplane_r);
  hipLaunchKernelGGL(k_plane_dot, dim3((unsigned)np), dim3(256), 0, b->stream, R0, R1,
                     b->psum + np, b->nplane_r);
  hipLaunchKernelGGL(k_bs_final, dim3(grid_1d(nr)), dim3(kGrid), 0, b->stream, R2, R1, b->psum,
                     b->psum + np, b->pHi, back_constant, eps, R3, b->M, b->nz,
                     (int)b->nplane_r);
  HIPC(hipGetLastError());
  if ((rc = param_emit(b, R3, b->field_r[QGK_DQ], dq_host, 1.0, weight, accumulate))) return rc;
  HIPC(stream_done(b));
  return 0;
}

}  // extern "C"
