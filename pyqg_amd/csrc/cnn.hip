// libqgcnn: BF16 MFMA kernels of the learned subgrid closure (gfx950).
//
// A fully convolutional network of 5x5 periodic ("circular") convolutions on
// square images, activations NHWC in BF16 with channels padded to multiples of
// 16, FP32 accumulation in the matrix cores (v_mfma_f32_16x16x16_bf16).
//
//   forward        y = act(conv(x, W) + b)            k_conv_fwd
//   backward data  dx = conv(dy, W') * [x_prev > 0]   k_conv_fwd on W' = flip/transpose(W)
//   backward wts   dW = sum_px dy (x) x_shifted        k_conv_wgrad
//   bias grad, MSE loss / grad, Adam, casts           small element-wise kernels
//
// Weight layout W[cout][25][cin] (tap = dy*5 + dx, dy, dx in 0..4 for offsets
// -2..2).  All shapes are checked on the host before any launch.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/pyqg_amd_cnn.h"

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

#define HIPC(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return fail(QGCNN_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));            \
  } while (0)

typedef unsigned short bf16_t;  // raw bit pattern
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t h) {
  return __uint_as_float(((unsigned)h) << 16);
}
// round to nearest even
__device__ __forceinline__ bf16_t f2bf(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (bf16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}

constexpr int KS = 5;       // kernel size
constexpr int NTAP = KS * KS;

// --------------------------------------------------------------------------
// Forward convolution.  Block: one output row y of sample b, a group of up to
// 64 output channels (one wave per 16), all W pixels (W/16 MFMA tiles per
// wave).  The five input rows y-2..y+2 of one 16-channel chunk are staged in
// LDS; every tap reads its shifted pixels from there.
// --------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void k_conv_fwd(const bf16_t* __restrict__ x,
                                                  const bf16_t* __restrict__ w,
                                                  const float* __restrict__ bias, int H,
                                                  int cinp, int coutp, int relu,
                                                  const bf16_t* __restrict__ mask,
                                                  bf16_t* __restrict__ y,
                                                  float* __restrict__ y32) {
  constexpr int NT = W / 16;  // pixel tiles per row
  __shared__ bf16_t lds[KS * W * 16];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nwaves = blockDim.x / 64;
  const int row = blockIdx.x, b = blockIdx.y;
  const int co0 = blockIdx.z * 64 + wave * 16;  // this wave's 16 output channels
  const bool active = co0 < coutp;
  v4f acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
  const int li = lane % 16, lk = lane / 16;
  for (int ci0 = 0; ci0 < cinp; ci0 += 16) {
    __syncthreads();
    // stage rows row-2..row+2 (periodic), W pixels, 16 channels: 16-byte pieces
    for (int q = threadIdx.x; q < KS * W * 2; q += blockDim.x) {
      const int half = q & 1, pix = (q >> 1) % W, r = (q >> 1) / W;
      const int yy = (row + r - 2 + H) % H;
      const uint4 v = *reinterpret_cast<const uint4*>(
          x + (((size_t)b * H + yy) * W + pix) * cinp + ci0 + half * 8);
      *reinterpret_cast<uint4*>(lds + (r * W + pix) * 16 + half * 8) = v;
    }
    __syncthreads();
    if (active) {
      for (int tap = 0; tap < NTAP; ++tap) {
        const int dy = tap / KS, dx = tap % KS;
        const v4s a = *reinterpret_cast<const v4s*>(
            w + ((size_t)(co0 + li) * NTAP + tap) * cinp + ci0 + lk * 4);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int pix = (t * 16 + li + dx - 2 + W) % W;
          const v4s bv = *reinterpret_cast<const v4s*>(lds + (dy * W + pix) * 16 + lk * 4);
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
        }
      }
    }
  }
  (void)nwaves;
  if (!active) return;
  // epilogue: lane holds channels co0 + 4*lk + j at pixel 16 t + li
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int pix = t * 16 + li;
    const size_t o = (((size_t)b * H + row) * W + pix) * coutp + co0 + lk * 4;
    float v[4] = {acc[t][0], acc[t][1], acc[t][2], acc[t][3]};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (bias) v[j] += bias[co0 + lk * 4 + j];
      if (relu) v[j] = v[j] > 0.f ? v[j] : 0.f;
    }
    if (mask) {
      const v4s m = *reinterpret_cast<const v4s*>(mask + o);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (bf2f((bf16_t)m[j]) <= 0.f) v[j] = 0.f;
    }
    if (y) {
      v4s out;
#pragma unroll
      for (int j = 0; j < 4; ++j) out[j] = (short)f2bf(v[j]);
      *reinterpret_cast<v4s*>(y + o) = out;
    }
    if (y32) {
#pragma unroll
      for (int j = 0; j < 4; ++j) y32[o + j] = v[j];
    }
  }
}

// --------------------------------------------------------------------------
// Weight gradient.  dW[co][tap][ci] += sum over the block's pixels of
// dy[px][co] * x[px shifted by tap][ci].  Block: one tap, 64 output channels
// (one wave per 16) x 16 input channels... x cinp/16 chunks, a range of rows of
// one sample.  Pixels are the reduction (K) dimension: 16-pixel slices of dy and
// x are transposed through LDS into the MFMA operand layout.
// --------------------------------------------------------------------------
template <int W>
__global__ __launch_bounds__(256) void k_conv_wgrad(const bf16_t* __restrict__ dy,
                                                    const bf16_t* __restrict__ x, int H,
                                                    int cinp, int coutp, int rows_per_block,
                                                    float* __restrict__ dw) {
  // LDS: dy tile [64 co][16 px] and x tile [cinp][16 px] (K contiguous)
  extern __shared__ bf16_t smem[];
  bf16_t* sdy = smem;              // 64 * 16
  bf16_t* sx = smem + 64 * 16;     // cinp * 16
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int tap = blockIdx.x % NTAP;
  const int cog = blockIdx.x / NTAP;          // 64-channel group of cout
  const int b = blockIdx.y;
  const int r0 = blockIdx.z * rows_per_block;
  const int dyo = tap / KS - 2, dxo = tap % KS - 2;
  const int co0 = cog * 64 + wave * 16;
  const bool active = co0 < coutp;
  const int nci = cinp / 16;
  const int li = lane % 16, lk = lane / 16;
  v4f acc[8];  // up to 8 ci tiles (cinp <= 128)
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
  const int cog_w = (coutp - cog * 64) < 64 ? (coutp - cog * 64) : 64;
  for (int r = r0; r < r0 + rows_per_block && r < H; ++r) {
    const int ys = (r + dyo + H) % H;
    for (int p0 = 0; p0 < W; p0 += 16) {
      __syncthreads();
      // dy tile: 16 px x cog_w channels -> sdy[c][p]
      for (int q = threadIdx.x; q < 16 * cog_w; q += blockDim.x) {
        const int c = q % cog_w, p = q / cog_w;
        sdy[c * 16 + p] = dy[(((size_t)b * H + r) * W + p0 + p) * coutp + cog * 64 + c];
      }
      // x tile (shifted): 16 px x cinp channels -> sx[c][p]
      for (int q = threadIdx.x; q < 16 * cinp; q += blockDim.x) {
        const int c = q % cinp, p = q / cinp;
        const int xs = (p0 + p + dxo + W) % W;
        sx[c * 16 + p] = x[(((size_t)b * H + ys) * W + xs) * cinp + c];
      }
      __syncthreads();
      if (active) {
        // A[i = co][k = px], B[k = px][n = ci]
        const v4s a = *reinterpret_cast<const v4s*>(sdy + (wave * 16 + li) * 16 + lk * 4);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          if (t < nci) {
            const v4s bv = *reinterpret_cast<const v4s*>(sx + (t * 16 + li) * 16 + lk * 4);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bv, acc[t], 0, 0, 0);
          }
        }
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    if (t >= nci) break;
    // lane holds D[i = co0 + 4 lk + j][n = ci = 16 t + li]
#pragma unroll
    for (int j = 0; j < 4; ++j)
      atomicAdd(dw + ((size_t)(co0 + lk * 4 + j) * NTAP + tap) * cinp + t * 16 + li, acc[t][j]);
  }
}

// W'[ci][24 - tap][co] = W[co][tap][ci]: the backward-data weights
__global__ void k_flip_transpose(const bf16_t* __restrict__ w, bf16_t* __restrict__ wt, int cinp,
                                 int coutp) {
  const size_t n = (size_t)coutp * NTAP * cinp;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int ci = (int)(i % cinp);
    const int tap = (int)((i / cinp) % NTAP);
    const int co = (int)(i / ((size_t)cinp * NTAP));
    wt[((size_t)ci * NTAP + (NTAP - 1 - tap)) * coutp + co] = w[i];
  }
}

// db[c] += sum over pixels of dy[px][c]
__global__ void k_bias_grad(const bf16_t* __restrict__ dy, size_t npix, int coutp,
                            float* __restrict__ db) {
  const int c = threadIdx.x % coutp;
  const int lanes_per_c = blockDim.x / coutp;
  const int sub = threadIdx.x / coutp;
  float s = 0.f;
  for (size_t p = blockIdx.x * (size_t)lanes_per_c + sub; p < npix;
       p += (size_t)gridDim.x * lanes_per_c)
    s += bf2f(dy[p * coutp + c]);
  if (sub < lanes_per_c) atomicAdd(db + c, s);
}

// MSE over the first `nch` of `coutp` channels: loss += sum (pred - target)^2 * inv_n;
// grad (bf16, all coutp channels, zero in the padding) = 2 (pred - target) * inv_n * gscale
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      size_t npix, int coutp, int nch, float inv_n, float gscale,
                      bf16_t* __restrict__ grad, float* __restrict__ loss) {
  float s = 0.f;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < npix * coutp;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % coutp);
    const size_t p = i / coutp;
    float g = 0.f;
    if (c < nch) {
      const float d = pred[i] - target[p * nch + c];
      s += d * d;
      g = 2.f * d * inv_n * gscale;
    }
    if (grad) grad[i] = f2bf(g);
  }
  // wave reduction then one atomic per wave
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, s * inv_n);
}

__global__ void k_f2bf(const float* __restrict__ a, bf16_t* __restrict__ o, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    o[i] = f2bf(a[i]);
}

__global__ void k_bf2f(const bf16_t* __restrict__ a, float* __restrict__ o, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    o[i] = bf2f(a[i]);
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, bf16_t* __restrict__ pb, size_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2, float wd) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * p[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] -= upd;
    if (pb) pb[i] = f2bf(p[i]);
  }
}

int grid1(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  return (int)(g ? g : 1);
}

bool mult16(int c) { return c > 0 && c % 16 == 0 && c <= 128; }

}  // namespace

extern "C" {

const char* qgcnn_last_error(void) { return g_err.c_str(); }

int qgcnn_conv_fwd(const void* x, int B, int H, int W, int cinp, const void* w,
                   const float* bias, int coutp, int relu, const void* mask, void* y, float* y32,
                   void* stream) {
  if (!x || !w || (!y && !y32)) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || !mult16(coutp))
    return fail(QGCNN_E_ARG, "channel counts must be multiples of 16 in [16, 128]");
  if (B < 1 || H < KS || W != 64)
    return fail(QGCNN_E_ARG, "images must be W=64 wide and at least 5 rows (got %dx%d)", H, W);
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(H, B, (coutp + 63) / 64);
  const int waves = coutp >= 64 ? 4 : coutp / 16;
  hipLaunchKernelGGL(k_conv_fwd<64>, grid, dim3(64 * waves), 0, st, (const bf16_t*)x,
                     (const bf16_t*)w, bias, H, cinp, coutp, relu, (const bf16_t*)mask,
                     (bf16_t*)y, y32);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_conv_wgrad(const void* dy, const void* x, int B, int H, int W, int cinp, int coutp,
                     float* dw, void* stream) {
  if (!dy || !x || !dw) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || !mult16(coutp))
    return fail(QGCNN_E_ARG, "channel counts must be multiples of 16 in [16, 128]");
  if (B < 1 || H < KS || W != 64) return fail(QGCNN_E_ARG, "images must be W=64 wide");
  hipStream_t st = (hipStream_t)stream;
  const int rows_per_block = 8;
  dim3 grid(NTAP * ((coutp + 63) / 64), B, (H + rows_per_block - 1) / rows_per_block);
  const int waves = coutp >= 64 ? 4 : coutp / 16;
  const size_t lds = (size_t)(64 * 16 + cinp * 16) * sizeof(bf16_t);
  hipLaunchKernelGGL(k_conv_wgrad<64>, grid, dim3(64 * waves), lds, st, (const bf16_t*)dy,
                     (const bf16_t*)x, H, cinp, coutp, rows_per_block, dw);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_flip_transpose(const void* w, void* wt, int cinp, int coutp, void* stream) {
  if (!w || !wt) return fail(QGCNN_E_ARG, "NULL argument");
  const size_t n = (size_t)coutp * NTAP * cinp;
  hipLaunchKernelGGL(k_flip_transpose, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream,
                     (const bf16_t*)w, (bf16_t*)wt, cinp, coutp);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_bias_grad(const void* dy, size_t npix, int coutp, float* db, void* stream) {
  if (!dy || !db || !mult16(coutp)) return fail(QGCNN_E_ARG, "bad arguments");
  hipLaunchKernelGGL(k_bias_grad, dim3(512), dim3(256 / coutp * coutp), 0, (hipStream_t)stream,
                     (const bf16_t*)dy, npix, coutp, db);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_mse(const float* pred, const float* target, size_t npix, int coutp, int nch,
              float gscale, void* grad, float* loss, void* stream) {
  if (!pred || !target || nch < 1 || nch > coutp) return fail(QGCNN_E_ARG, "bad arguments");
  const float inv_n = 1.0f / (float)((double)npix * nch);
  hipLaunchKernelGGL(k_mse, dim3(grid1(npix * coutp)), dim3(256), 0, (hipStream_t)stream, pred,
                     target, npix, coutp, nch, inv_n, gscale, (bf16_t*)grad, loss);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_f32_to_bf16(const float* a, void* o, size_t n, void* stream) {
  hipLaunchKernelGGL(k_f2bf, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, a, (bf16_t*)o, n);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_bf16_to_f32(const void* a, float* o, size_t n, void* stream) {
  hipLaunchKernelGGL(k_bf2f, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream,
                     (const bf16_t*)a, o, n);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_adam(float* p, const float* g, float* m, float* v, void* pb, size_t n, float lr,
               float b1, float b2, float eps, int step, float wd, void* stream) {
  if (!p || !g || !m || !v || step < 1) return fail(QGCNN_E_ARG, "bad arguments");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, (hipStream_t)stream, p, g, m, v,
                     (bf16_t*)pb, n, lr, b1, b2, eps, c1, c2, wd);
  HIPC(hipGetLastError());
  return 0;
}

}  // extern "C"
