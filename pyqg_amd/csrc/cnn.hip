// libqgcnn: BF16 MFMA kernels of the learned subgrid closure (gfx950).
//
// A fully convolutional network of 5x5 periodic ("circular") convolutions on
// square images, activations NHWC in BF16 with channels padded to multiples of
// 16, FP32 accumulation in the matrix cores (v_mfma_f32_16x16x32_bf16, and the
// K=16 form for 16-channel inputs).
//
//   forward        y = act(conv(x, W) + b)            k_conv_fwd_ws (persistent, overlapped
//                                                      staging; 64-pixel rows, heights a
//                                                      multiple of 8), k_conv_fwd_mt else;
//                                                      first / last layers k_conv_first_ws,
//                                                      k_conv_last_ws (k_conv_first/_last)
//   backward data  dx = conv(dy, W') * [x_prev > 0]   the forward kernels on W' = flip/transpose(W)
//   backward wts   dW = sum_px dy (x) x_shifted        k_conv_wgrad_mt (+ k_wgrad_reduce),
//                                                      or k_conv_wgrad with atomics
//   bias grad, MSE loss / grad, Adam, casts           small element-wise kernels
//
// Weight layout W[cout][25][cin] (tap = dy*5 + dx, dy, dx in 0..4 for offsets
// -2..2).  All shapes are checked on the host before any launch.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <hip/hip_bf16.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <initializer_list>
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

// compute units of the current device (queried once per device)
int num_cus() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n < 1)
      n = 256;
    cache[dev] = n;
  }
  return cache[dev];
}

constexpr int kLoaderWaves = 4;  // staging waves of k_conv_fwd_ws

typedef unsigned short bf16_t;  // raw bit pattern
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(bf16_t h) {
  return __uint_as_float(((unsigned)h) << 16);
}
// round to nearest even: the compiler's float -> bfloat conversion, one
// v_cvt_pk_bf16_f32 on gfx950 (NaN stays NaN, quieted)
__device__ __forceinline__ bf16_t f2bf(float f) {
  const __bf16 h = (__bf16)f;
  return __builtin_bit_cast(bf16_t, h);
}

constexpr int KS = 5;       // kernel size
constexpr int NTAP = KS * KS;

// --------------------------------------------------------------------------
// Forward convolution, blocked for the gfx950 matrix cores.  A block of four
// waves computes 4*RW consecutive output rows (RW per wave) of one sample for
// CO = 16*MT output channels and all 64 pixels: MT x 4*RW MFMA tiles per wave.
// The K loop runs over chunks of KC input channels (KC = 32 uses the K=32
// instruction v_mfma_f32_16x16x32_bf16, KC = 16 the K=16 one); per chunk the
// 4*RW + 4 input rows the block's output rows need are staged in LDS once, with a
// padded pixel stride against bank conflicts.  Every B operand read from LDS
// feeds MT MFMAs; the A operands (weights) come from the L1/L2-resident tensor.
// MFMA lane layout (A 16 x K, B K x 16): lane (i = lane%16, g = lane/16) holds
// A[i][g*K/4 .. +K/4] and B[g*K/4 .. +K/4][i]; the accumulator lane holds
// D[4g + j][i].  A and B use the same K slots, so the sum covers every channel.
// Row r of tile m is output channel tile_channel<MT>(co0, m, r) (below), so a
// lane's accumulators over the MT tiles are consecutive channels of its pixel.
// --------------------------------------------------------------------------
// compile-time loop: f(std::integral_constant<int, I>{}) for I in [B, E)
template <int I, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < E) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, E>(f);
  }
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

template <int KC>
struct MfmaK;
template <>
struct MfmaK<32> {
  typedef v8bf op;
  __device__ __forceinline__ static v4f mma(op a, op b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct MfmaK<16> {
  typedef v4s op;
  __device__ __forceinline__ static v4f mma(op a, op b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
  }
};

// Output channels of a block's MT matrix tiles.  The rows of the tiles are
// assigned so that accumulator lane group g (rows 4g..4g+3 of every tile)
// holds the 4*MT consecutive channels co0 + 4*MT*g + [0, 4*MT): tile m, row r
// computes channel co0 + 4*MT*(r/4) + 4*m + r%4.  A lane then stores each of
// its pixels' channels as one contiguous 8*MT-byte run, and the four lane
// groups together write the block's 32*MT bytes of the pixel in one piece
// (one 128-byte line at MT = 4) instead of MT scattered 32-byte fragments.
template <int MT>
__device__ __forceinline__ int tile_channel(int co0, int m, int r) {
  return co0 + 4 * MT * (r / 4) + 4 * m + r % 4;
}

typedef short v8s __attribute__((ext_vector_type(8)));

// Epilogue of one pixel's 4*MT consecutive channels (flat NHWC offset o):
// bias (bv, read once per lane by load_bias), ReLU, mask (channels whose mask
// value is <= 0 are zeroed), then BF16 and/or FP32 stores, each as whole
// 16-byte pieces where MT >= 2.
template <int MT>
__device__ __forceinline__ void load_bias(const float* __restrict__ bias, int cb,
                                          float (&bv)[4 * MT]) {
#pragma unroll
  for (int p = 0; p < MT; ++p) {
    const v4f q = bias ? *reinterpret_cast<const v4f*>(bias + cb + 4 * p) : v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; ++j) bv[4 * p + j] = q[j];
  }
}

template <int MT>
__device__ __forceinline__ void store_channels(float (&v)[4 * MT], const float (&bv)[4 * MT],
                                               bool has_bias, int relu,
                                               const bf16_t* __restrict__ mask,
                                               bf16_t* __restrict__ y, float* __restrict__ y32,
                                               size_t o) {
#pragma unroll
  for (int i = 0; i < 4 * MT; ++i) {
    if (has_bias) v[i] += bv[i];
    if (relu) v[i] = v[i] > 0.f ? v[i] : 0.f;
  }
  if (mask) {
#pragma unroll
    for (int p = 0; p < MT; ++p) {
      const v4s mk = *reinterpret_cast<const v4s*>(mask + o + 4 * p);
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (bf2f((bf16_t)mk[j]) <= 0.f) v[4 * p + j] = 0.f;
    }
  }
  if (y) {
    if constexpr (MT >= 2) {
#pragma unroll
      for (int p = 0; p < MT / 2; ++p) {
        v8s out;
#pragma unroll
        for (int j = 0; j < 8; ++j) out[j] = (short)f2bf(v[8 * p + j]);
        *reinterpret_cast<v8s*>(y + o + 8 * p) = out;
      }
    } else {
      v4s out;
#pragma unroll
      for (int j = 0; j < 4; ++j) out[j] = (short)f2bf(v[j]);
      *reinterpret_cast<v4s*>(y + o) = out;
    }
  }
  if (y32) {
#pragma unroll
    for (int p = 0; p < MT; ++p)
      *reinterpret_cast<v4f*>(y32 + o + 4 * p) = v4f{v[4 * p], v[4 * p + 1], v[4 * p + 2], v[4 * p + 3]};
  }
}

// Staging of input rows row0 - 2 .. row0 - 2 + SR - 1 (periodic in y) of sample
// b, channels ci0 .. ci0 + 8*PIECES - 1, into lds[(r*W + pix)*PS + c] by the
// block's 256 threads.  Every thread issues all of its global loads before its
// first LDS write, so the block waits for one memory latency per staging
// instead of one per 16-byte piece.
template <int SR, int W, int PIECES, int PS>
__device__ __forceinline__ void stage_rows(const bf16_t* __restrict__ x, bf16_t* lds, int b,
                                           int H, int row0, int cinp, int ci0) {
  constexpr int TOTAL = SR * W * PIECES;
  constexpr int NQ = (TOTAL + 255) / 256;
  uint4 v[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int q = threadIdx.x + 256 * i;
    if (TOTAL % 256 == 0 || q < TOTAL) {
      const int pc = q % PIECES, pix = (q / PIECES) % W, r = q / (PIECES * W);
      const int yy = ((row0 + r - 2) % H + H) % H;
      v[i] = *reinterpret_cast<const uint4*>(x + (((size_t)b * H + yy) * W + pix) * cinp + ci0 +
                                             pc * 8);
    }
  }
#pragma unroll
  for (int i = 0; i < NQ; ++i) {
    const int q = threadIdx.x + 256 * i;
    if (TOTAL % 256 == 0 || q < TOTAL) {
      const int pc = q % PIECES, pix = (q / PIECES) % W, r = q / (PIECES * W);
      *reinterpret_cast<uint4*>(lds + (r * W + pix) * PS + pc * 8) = v[i];
    }
  }
}

template <int W, int MT, int KC, int RW>
__global__ __launch_bounds__(256, RW) void k_conv_fwd_mt(const bf16_t* __restrict__ x,
                                                     const bf16_t* __restrict__ w,
                                                     const float* __restrict__ bias, int H,
                                                     int cinp, int coutp, int relu,
                                                     const bf16_t* __restrict__ mask,
                                                     bf16_t* __restrict__ y,
                                                     float* __restrict__ y32) {
  using M = MfmaK<KC>;
  typedef typename M::op op;
  constexpr int NT = W / 16;            // pixel tiles per row
  constexpr int ROWS = 4 * RW;          // output rows per block, RW per wave
  constexpr int SR = ROWS + KS - 1;     // staged input rows
  constexpr int PS = KC + 8;            // padded pixel stride (bf16 elements)
  constexpr int KG = KC / 4;            // K elements per lane
  constexpr int PIECES = KC / 8;        // 16-byte pieces per staged pixel
  __shared__ __attribute__((aligned(16))) bf16_t lds[SR * W * PS];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int row0 = blockIdx.x * ROWS, b = blockIdx.y;
  const int co0 = blockIdx.z * (16 * MT);
  const int wrow0 = row0 + wave * RW;   // the wave's first output row
  const bool wave_ok = wrow0 < H;
  const int li = lane % 16, lg = lane / 16;
  v4f acc[MT][RW][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ci0 = 0; ci0 < cinp; ci0 += KC) {
    __syncthreads();
    stage_rows<SR, W, PIECES, PS>(x, lds, b, H, row0, cinp, ci0);
    __syncthreads();
    if (wave_ok) {
      const bf16_t* wrow =
          w + (size_t)tile_channel<MT>(co0, 0, li) * NTAP * cinp + ci0 + lg * KG;
      auto load_a = [&](op (&a)[MT], int tap) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
          a[m] = *reinterpret_cast<const op*>(wrow + ((size_t)m * 4 * NTAP + tap) * cinp);
      };
      // One tap: every B operand read from LDS before the first matrix op (the
      // LDS latency is paid once per tap, not per operand pair), then the
      // MT x RW x NT matrix ops.  Up to two weight tiles, the weight operands
      // of the next tap are loaded first, so their L2 latency overlaps this
      // tap's work (two register sets, taps taken in pairs).  The scheduling
      // barriers keep
      // that order: the compiler's own schedule interleaves the reads with
      // the matrix ops two at a time and waits for each pair.
      auto tap_step = [&](const op (&a)[MT], op (&a_next)[MT], int tap, auto next_c) {
        constexpr bool next = decltype(next_c)::value;
        if constexpr (next) load_a(a_next, tap + 1);
        const int dy = tap / KS, dx = tap % KS;
        op bv[RW][NT];
#pragma unroll
        for (int rr = 0; rr < RW; ++rr) {
          const bf16_t* lrow = lds + (wave * RW + rr + dy) * W * PS + lg * KG;
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int pix = (t * 16 + li + dx - 2 + W) % W;
            bv[rr][t] = *reinterpret_cast<const op*>(lrow + pix * PS);
          }
        }
#pragma unroll
        for (int rr = 0; rr < RW; ++rr)
#pragma unroll
          for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m) acc[m][rr][t] = M::mma(a[m], bv[rr][t], acc[m][rr][t]);
        __builtin_amdgcn_sched_group_barrier(0x020, next ? MT : 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, RW * NT, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, MT * RW * NT, 0);
      };
      op a0[MT], a1[MT];
      if constexpr (MT <= 2) {
        load_a(a0, 0);
#pragma unroll 1
        for (int tap = 0; tap < NTAP - 1; tap += 2) {  // NTAP is odd
          tap_step(a0, a1, tap, std::true_type{});
          tap_step(a1, a0, tap + 1, std::true_type{});
        }
        tap_step(a0, a1, NTAP - 1, std::false_type{});
      } else {  // four weight tiles: no room for a second register set
#pragma unroll 1
        for (int tap = 0; tap < NTAP; ++tap) {
          load_a(a0, tap);
          tap_step(a0, a1, tap, std::false_type{});
        }
      }
    }
  }
  if (!wave_ok) return;
  const int cb = tile_channel<MT>(co0, 0, 4 * lg);
  float bv[4 * MT];
  load_bias<MT>(bias, cb, bv);
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wrow0 + rr;
    if (row >= H) break;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const size_t o = (((size_t)b * H + row) * W + t * 16 + li) * coutp + cb;
      float v[4 * MT];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[4 * m + j] = acc[m][rr][t][j];
      store_channels<MT>(v, bv, bias != nullptr, relu, mask, y, y32, o);
    }
  }
}

// --------------------------------------------------------------------------
// Forward convolution with overlapped staging (the K=32 form, 64-pixel rows,
// heights a multiple of 8, 32- or 64-channel output groups).  k_conv_fwd_mt
// stages a chunk's input rows, waits, and only then runs the chunk's matrix
// ops, so each block alternates between a memory phase and a compute phase.
// Here the blocks stay resident for the whole layer, each walking its share of
// the work items (sample, 8-row group, output-channel group) as cinp/32
// stages.  Warp roles:
//   waves 0-3 (compute): the matrix ops on the stage in LDS, the next tap's
//     weight operands and the next row's LDS operands loaded while the current
//     ones feed the matrix cores.  With four weight tiles (64 channels) the
//     waves pair up, two tiles x four rows each; with two, each wave takes
//     both tiles for two rows;
//   waves 4..4+NLW-1 (loaders): read the next stage's input rows from HBM.
// DB (four weight tiles: one block per CU, two LDS buffers): the loaders write
// stage s into buffer s%2 and arrive at barrier s; the compute waves pass
// barrier s and then read buffer s%2.  A loader's write of stage s+1 (after
// barrier s) goes to buffer (s+1)%2, whose last readers (stage s-1) all
// arrived at barrier s before that.  One buffer (two weight tiles: two blocks
// per CU, so that each SIMD also interleaves two compute waves): the loaders
// hold the next stage in registers; per stage, barrier A (the compute waves
// are done with the previous stage), the write, barrier B, then the loads of
// the following stage, which overlap this stage's matrix ops.
// The roles have separate memory counters, so a compute wave's weight loads
// never wait for the staging loads.
// Measured at 128 -> 64 channels, batch 128 (variants timed outside the
// library): the kernel ~210 us (k_conv_fwd_mt ~255 us); 43 us with the matrix
// ops removed; 134 us with the staging removed (LDS never written); ~170 us
// with pseudo-random data written to LDS each stage and no HBM loads.  The
// matrix ops run slower on operand data that toggles (a matrix-op-only loop:
// 1.39 PFLOP/s on random operands, 1.56 on constant ones) and slower again
// beside the HBM traffic of the staging.  At 32 -> 32 channels: 30 us against
// 33 for k_conv_fwd_mt (the DB form with one compute wave per SIMD: 37-42).
// --------------------------------------------------------------------------
template <int W, int MT, int RW, int NLW, bool DB>
__global__ __launch_bounds__(256 + 64 * NLW, DB ? 1 : 2) void k_conv_fwd_ws(
    const bf16_t* __restrict__ x, const bf16_t* __restrict__ w, const float* __restrict__ bias,
    int B, int H, int cinp, int coutp, int relu, const bf16_t* __restrict__ mask,
    bf16_t* __restrict__ y, float* __restrict__ y32) {
  constexpr int KC = 32;
  using M = MfmaK<KC>;
  typedef typename M::op op;
  constexpr int NT = W / 16;
  constexpr int ROWS = 4 * RW;
  constexpr int SR = ROWS + KS - 1;
  constexpr int KG = KC / 4;
  constexpr int PIECES = KC / 8;
  // LDS stage layout [row][K group][pixel - 2 .. pixel W + 1], 16 bytes per
  // entry, the row's periodic halo stored twice: the 16 lanes of an operand
  // read that share a K group read 16 consecutive entries, i.e. 256 contiguous
  // bytes, free of bank conflicts for every tap's shift (a pixel-major layout
  // with 64- or 80-byte pixels delivers about half the LDS bandwidth for this
  // pattern), and a read's address is a per-lane base plus constants (no
  // per-read wrap-around arithmetic)
  constexpr int WP = W + 4;
  constexpr int BUF = SR * PIECES * WP * 8;
  __shared__ __attribute__((aligned(16))) bf16_t lds[(DB ? 2 : 1) * BUF];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nrg = H / ROWS, nch = cinp / KC, ncog = coutp / (16 * MT);
  const int items = B * nrg * ncog;
  const int mine = (int)blockIdx.x < items ? (items - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
  const int S = mine * nch;

  if (wave >= 4) {  // loader waves
    // thread lt stages K group pc of pixel pix in every staged row: a wave
    // covers 16 pixels x 4 K groups (1 KB of global memory) and writes 16
    // consecutive entries of each K group's plane.  The row offsets are
    // wave-uniform (scalar registers); the lane's part is fixed.
    // With DB, stage s is loaded and written to buffer s%2, then barrier s.
    // Without (one buffer, two blocks per CU): the loaders hold the next
    // stage in registers; per stage, barrier A (the compute waves are done
    // with the previous stage), the write, barrier B, then the loads of the
    // stage after it, which overlap this stage's matrix ops.
    constexpr int NL = 64 * NLW;
    constexpr int K = PIECES * W / NL;  // loader passes per staged row
    static_assert(K * NL == PIECES * W, "whole rows per loader passes");
    const int lt = threadIdx.x - 256;
    const __amdgpu_buffer_rsrc_t xrs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)x, (short)0, B * H * W * cinp * (int)sizeof(bf16_t), 0x00020000);
    int xlane[K], halo[K];
    bf16_t* lw[K];
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const int qq = lt + NL * j;
      const int pix = ((qq / 64) % (W / 16)) * 16 + qq % 16, pc = (qq / 16) % PIECES;
      xlane[j] = (pix * cinp + pc * 8) * (int)sizeof(bf16_t);
      lw[j] = lds + (pc * WP + pix + 2) * 8;
      // the halo copies: pixels W-2, W-1 also left of pixel 0, pixels 0, 1
      // also right of pixel W-1
      halo[j] = pix >= W - 2 ? -W : (pix < 2 ? W : 0);
    }
    uint4 v[SR * K];
    auto load = [&](int st) {
      const int it = blockIdx.x + (st / nch) * gridDim.x;
      const int ci0 = (st % nch) * KC;
      const int rest = it / ncog;
      const int b = rest / nrg, row0 = (rest % nrg) * ROWS;
#pragma unroll
      for (int r = 0; r < SR; ++r) {
        const int yy = (row0 + r - 2 + H) % H;
        const int so = ((b * H + yy) * W * cinp + ci0) * (int)sizeof(bf16_t);
#pragma unroll
        for (int j = 0; j < K; ++j)
          v[r * K + j] = __builtin_bit_cast(
              uint4, __builtin_amdgcn_raw_buffer_load_b128(xrs, xlane[j], so, 0));
      }
    };
    auto write = [&](int boff) {
#pragma unroll
      for (int j = 0; j < K; ++j) {
        bf16_t* dst = lw[j] + boff;
#pragma unroll
        for (int r = 0; r < SR; ++r)
          *reinterpret_cast<uint4*>(dst + r * PIECES * WP * 8) = v[r * K + j];
        if (halo[j]) {
#pragma unroll
          for (int r = 0; r < SR; ++r)
            *reinterpret_cast<uint4*>(dst + (r * PIECES * WP + halo[j]) * 8) = v[r * K + j];
        }
      }
    };
    if constexpr (DB) {
      for (int s = 0; s < S; ++s) {
        load(s);
        write((s & 1) * BUF);
        __syncthreads();
      }
    } else {
      if (S > 0) load(0);
      for (int s = 0; s < S; ++s) {
        __syncthreads();  // A
        write(0);
        __syncthreads();  // B
        if (s + 1 < S) load(s + 1);
      }
    }
    return;
  }

  const int li = lane % 16, lg = lane / 16;
  // Wave tiles: the waves pair up over the four weight tiles (64 output
  // channels), each wave taking two of the tiles for four of the block's eight
  // rows, so a weight operand is loaded by two waves instead of four (with
  // every wave loading all four, the weight loads, 16 lines each, kept the
  // address unit stalled on the L1 for ~45 % of the kernel) and each LDS
  // operand feeds two matrix ops.
  // (with two weight tiles each wave takes both for RW rows)
  static_assert(RW == 2, "the LDS operand ring alternates between the wave's rows");
  constexpr bool PAIR = MT >= 4;
  constexpr int MTW = PAIR ? MT / 2 : MT;  // weight tiles per wave
  constexpr int RWW = PAIR ? 2 * RW : RW;  // output rows per wave
  const int mo = PAIR ? (wave & 1) * MTW : 0;             // the wave's first tile
  const int wr0 = PAIR ? (wave >> 1) * RWW : wave * RWW;  // the wave's first row
  // Operand pipeline of a compute wave, one step per (tap, output row): the
  // weights are loaded PD taps ahead (a ring of PD + 1 register sets; an L2
  // hit takes longer than one tap's matrix ops; the last taps of a stage load
  // the first taps of the block's next stage) and the LDS operands of the next
  // step one step ahead (two sets of one row's NT operands), so that one
  // resident compute wave per SIMD keeps its matrix ops fed without waiting
  // for either memory.
  constexpr int PD = 1;
  constexpr int RING = PD + 1;
  static_assert((NTAP - 1) % RING == 0, "the tap loop runs whole rings");
  const __amdgpu_buffer_rsrc_t wrs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)w, (short)0, coutp * NTAP * cinp * (int)sizeof(bf16_t), 0x00020000);
  const int wlane = (tile_channel<MT>(0, mo, li) * NTAP * cinp + lg * KG) * (int)sizeof(bf16_t);
  auto stage_w = [&](int st) {  // byte offset of the stage's weights
    const int it = blockIdx.x + (st / nch) * gridDim.x;
    const int co0 = (it % ncog) * 16 * MT;
    return (co0 * NTAP * cinp + (st % nch) * KC) * (int)sizeof(bf16_t);
  };
  auto load_a = [&](op (&a)[MTW], int wso, int tap) {
#pragma unroll
    for (int m = 0; m < MTW; ++m)
      a[m] = __builtin_bit_cast(
          op, __builtin_amdgcn_raw_buffer_load_b128(
                  wrs, wlane, wso + (m * 4 * NTAP + tap) * cinp * (int)sizeof(bf16_t), 0));
  };
  v4f acc[MTW][RWW][NT];
#pragma unroll
  for (int m = 0; m < MTW; ++m)
#pragma unroll
    for (int rr = 0; rr < RWW; ++rr)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
  op a[RING][MTW];
  op bq[2][NT];
  if (S > 0) {
    const int w0 = stage_w(0);
#pragma unroll
    for (int j = 0; j < PD; ++j) load_a(a[j], w0, j);
  }
  for (int s = 0; s < S; ++s) {
    __syncthreads();
    if constexpr (!DB) __syncthreads();  // B: the loaders' write of stage s is done
    const int it = blockIdx.x + (s / nch) * gridDim.x;
    const int ch = s % nch;
    const int cog = it % ncog, rest = it / ncog;
    const int b = rest / nrg, row0 = (rest % nrg) * ROWS;
    const int co0 = cog * 16 * MT;
    const bf16_t* buf = lds + (DB ? (s & 1) * BUF : 0);
    const int wr = stage_w(s);
    const bool more = s + 1 < S;
    const int wn = more ? stage_w(s + 1) : wr;
    // lane base: entry (row wr0, K group lg, pixel li - 2) of this stage's buffer
    const bf16_t* lbase = buf + ((wr0 * PIECES + lg) * WP + li) * 8;
    auto read_row = [&](op (&bv)[NT], int tap, int rr) {
      const int dy = tap / KS, dx = tap % KS;
      const bf16_t* lrow = lbase + ((rr + dy) * PIECES * WP + dx) * 8;
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const op*>(lrow + t * 16 * 8);
    };
    // one step: (tap, row RR) with weight set P = tap % RING; TAIL marks the
    // stage's last tap
    auto step = [&](int tap, auto pc, auto rrc, auto tailc) {
      constexpr int P = decltype(pc)::value;
      constexpr int RR = decltype(rrc)::value;
      constexpr bool TAIL = decltype(tailc)::value;
      constexpr bool lda = RR == 0;                  // weights PD taps on
      constexpr bool ldb = !(TAIL && RR == RWW - 1);  // LDS operands of the next step
      if constexpr (lda) {
        const int ta = tap + PD;
        if (ta < NTAP) load_a(a[(P + PD) % RING], wr, ta);
        else if (more) load_a(a[(P + PD) % RING], wn, ta - NTAP);
      }
      if constexpr (ldb) {
        if constexpr (RR + 1 < RWW) read_row(bq[(RR + 1) & 1], tap, RR + 1);
        else read_row(bq[0], tap + 1, 0);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < MTW; ++m)
          acc[m][RR][t] = M::mma(a[P][m], bq[RR & 1][t], acc[m][RR][t]);
      __builtin_amdgcn_sched_group_barrier(0x020, lda ? MTW : 0, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, ldb ? NT : 0, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, MTW * NT, 0);
    };
    read_row(bq[0], 0, 0);
#pragma unroll 1
    for (int tap = 0; tap < NTAP - 1; tap += RING) {
      static_for<0, RING>([&](auto pc) {
        static_for<0, RWW>([&](auto rc) { step(tap + decltype(pc)::value, pc, rc, std::false_type{}); });
      });
    }
    static_for<0, RWW>([&](auto rc) {
      step(NTAP - 1, std::integral_constant<int, (NTAP - 1) % RING>{}, rc, std::true_type{});
    });
    // the next stage's tap j (j < PD) was loaded into set (NTAP + j) % RING
    {
      op t[PD][MTW];
#pragma unroll
      for (int j = 0; j < PD; ++j)
#pragma unroll
        for (int m = 0; m < MTW; ++m) t[j][m] = a[(NTAP + j) % RING][m];
#pragma unroll
      for (int j = 0; j < PD; ++j)
#pragma unroll
        for (int m = 0; m < MTW; ++m) a[j][m] = t[j][m];
    }
    if (ch == nch - 1) {
      const int cb = tile_channel<MT>(co0, mo, 4 * lg);
      float bvv[4 * MTW];
      load_bias<MTW>(bias, cb, bvv);
#pragma unroll
      for (int rr = 0; rr < RWW; ++rr) {
        const int row = row0 + wr0 + rr;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const size_t o = (((size_t)b * H + row) * W + t * 16 + li) * coutp + cb;
          float v[4 * MTW];
#pragma unroll
          for (int m = 0; m < MTW; ++m)
#pragma unroll
            for (int j = 0; j < 4; ++j) v[4 * m + j] = acc[m][rr][t][j];
          store_channels<MTW>(v, bvv, bias != nullptr, relu, mask, y, y32, o);
#pragma unroll
          for (int m = 0; m < MTW; ++m) acc[m][rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
  }
}

// --------------------------------------------------------------------------
// First layer (at most 8 real input channels): the taps go into the K
// dimension.  K = tap * 8 + channel over the first 8 channels of the stored
// input and 25 taps padded to 28, so one K=32 MFMA covers 4 taps and the layer
// takes 7 matrix steps per tile instead of 25 K=16 ones over 16 padded
// channels.  In the operand layout lane group g holds K slots [8g, 8g + 8):
// tap 4c + g of step c, i.e. one 16-byte read of the staged pixel shifted by
// that tap.  Weights are repacked once per update: w8[co][28][8].
// --------------------------------------------------------------------------
constexpr int kFirstCh = 8;
constexpr int kFirstTaps = 28;

template <int W, int MT, int RW>
__global__ __launch_bounds__(256, RW) void k_conv_first(const bf16_t* __restrict__ x,
                                                       const bf16_t* __restrict__ w8,
                                                       const float* __restrict__ bias, int H,
                                                       int cinp, int coutp, int relu,
                                                       bf16_t* __restrict__ y,
                                                       float* __restrict__ y32) {
  typedef v8bf op;
  constexpr int NT = W / 16;
  constexpr int ROWS = 4 * RW;
  constexpr int SR = ROWS + KS - 1;
  constexpr int CH = kFirstCh;
  __shared__ __attribute__((aligned(16))) bf16_t lds[SR * W * CH];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int row0 = blockIdx.x * ROWS, b = blockIdx.y;
  const int co0 = blockIdx.z * (16 * MT);
  const int wrow0 = row0 + wave * RW;
  const bool wave_ok = wrow0 < H;
  const int li = lane % 16, lg = lane / 16;
  v4f acc[MT][RW][NT];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int rr = 0; rr < RW; ++rr)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
  stage_rows<SR, W, 1, CH>(x, lds, b, H, row0, cinp, 0);
  __syncthreads();
  if (wave_ok) {
    const bf16_t* wrow = w8 + (size_t)tile_channel<MT>(co0, 0, li) * kFirstTaps * CH + lg * CH;
#pragma unroll
    for (int c = 0; c < kFirstTaps / 4; ++c) {
      // taps 25-27 have zero weights; their operand reads stay on tap 24
      const int tap = 4 * c + lg < NTAP ? 4 * c + lg : NTAP - 1;
      const int dy = tap / KS, dx = tap % KS;
      op a[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m)
        a[m] = *reinterpret_cast<const op*>(wrow + ((size_t)m * 4 * kFirstTaps + 4 * c) * CH);
#pragma unroll
      for (int rr = 0; rr < RW; ++rr) {
        const bf16_t* lrow = lds + (wave * RW + rr + dy) * W * CH;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int pix = (t * 16 + li + dx - 2 + W) % W;
          const op bv = *reinterpret_cast<const op*>(lrow + pix * CH);
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][rr][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m], bv, acc[m][rr][t], 0, 0, 0);
        }
      }
    }
  }
  if (!wave_ok) return;
  const int cb = tile_channel<MT>(co0, 0, 4 * lg);
  float bv[4 * MT];
  load_bias<MT>(bias, cb, bv);
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wrow0 + rr;
    if (row >= H) break;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const size_t o = (((size_t)b * H + row) * W + t * 16 + li) * coutp + cb;
      float v[4 * MT];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int j = 0; j < 4; ++j) v[4 * m + j] = acc[m][rr][t][j];
      store_channels<MT>(v, bv, bias != nullptr, relu, nullptr, y, y32, o);
    }
  }
}

// First layer with overlapped staging (64-output-channel groups, 64-pixel
// rows, heights a multiple of 8): one resident block per CU walks its share of
// the (sample, 8-row group) items of one output-channel group, whose packed
// weights (4 tiles x 7 matrix steps x 64 lanes x 16 bytes = 28 KB) stay in LDS
// for the whole layer.  Four loader waves stage the next item's 12 input rows
// (8 channels, the periodic halo stored twice: [row][pixel - 2 .. 65]) into one
// of two LDS buffers while the four compute waves run the current item and
// store its outputs; the compute waves load nothing from global memory, so
// their stores drain while the next item's matrix ops run.  Operands are
// read one (step, row) ahead from LDS.  K layout as k_conv_first.
template <int W, int NLW>
__global__ __launch_bounds__(256 + 64 * NLW, 1) void k_conv_first_ws(
    const bf16_t* __restrict__ x, const bf16_t* __restrict__ w8, const float* __restrict__ bias,
    int B, int H, int cinp, int coutp, int relu, bf16_t* __restrict__ y,
    float* __restrict__ y32) {
  typedef v8bf op;
  constexpr int MT = 4, RW = 2, NT = W / 16;
  constexpr int ROWS = 4 * RW, SR = ROWS + KS - 1, CH = kFirstCh;
  constexpr int WP = W + 4;
  constexpr int BUF = SR * WP * CH;                // one input stage (bf16 elements)
  constexpr int NSTEP = kFirstTaps / 4;            // K=32 matrix steps
  constexpr int WSZ = MT * NSTEP * 64 * 8;         // the group's packed weights
  __shared__ __attribute__((aligned(16))) bf16_t lds[WSZ + 2 * BUF];
  bf16_t* const lw = lds;
  bf16_t* const lin = lds + WSZ;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nrg = H / ROWS, ncog = coutp / (16 * MT);
  const int items = B * nrg * ncog;
  // the grid is a multiple of ncog, so every item of a block has its group
  const int cog = (int)blockIdx.x % ncog, co0 = cog * 16 * MT;
  const int mine = (int)blockIdx.x < items ? (items - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
  {  // the group's weights in operand order: entry ((m * NSTEP + c) * 64 + lane)
    for (int e = threadIdx.x; e < MT * NSTEP * 64; e += blockDim.x) {
      const int ln = e % 64, c = (e / 64) % NSTEP, m = e / (64 * NSTEP);
      const int li = ln % 16, lg = ln / 16;
      *reinterpret_cast<uint4*>(lw + e * 8) = *reinterpret_cast<const uint4*>(
          w8 + ((size_t)tile_channel<MT>(co0, m, li) * kFirstTaps + 4 * c + lg) * CH);
    }
  }
  if (wave >= 4) {  // loader waves: stage of item k into buffer k%2, then barrier k
    constexpr int NL = 64 * NLW;
    constexpr int TOTAL = SR * W;
    constexpr int NQ = (TOTAL + NL - 1) / NL;
    const int lt = threadIdx.x - 256;
    for (int k = 0; k < mine; ++k) {
      const int it = blockIdx.x + k * gridDim.x;
      const int rest = it / ncog;
      const int b = rest / nrg, row0 = (rest % nrg) * ROWS;
      uint4 v[NQ];
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const int q = lt + NL * i;
        if (TOTAL % NL == 0 || q < TOTAL) {
          const int pix = q % W, r = q / W;
          const int yy = (row0 + r - 2 + H) % H;
          v[i] = *reinterpret_cast<const uint4*>(x + (((size_t)b * H + yy) * W + pix) * cinp);
        }
      }
      bf16_t* buf = lin + (k & 1) * BUF;
#pragma unroll
      for (int i = 0; i < NQ; ++i) {
        const int q = lt + NL * i;
        if (TOTAL % NL == 0 || q < TOTAL) {
          const int pix = q % W, r = q / W;
          bf16_t* dst = buf + (r * WP + pix + 2) * CH;
          *reinterpret_cast<uint4*>(dst) = v[i];
          const int halo = pix >= W - 2 ? -W : (pix < 2 ? W : 0);
          if (halo) *reinterpret_cast<uint4*>(dst + halo * CH) = v[i];
        }
      }
      __syncthreads();
    }
    return;
  }
  const int li = lane % 16, lg = lane / 16;
  const int wr0 = wave * RW;
  // per matrix step c the lane's tap 4c + lg (taps 25-27 have zero weights;
  // their operand reads stay on tap 24)
  int boff[NSTEP];
#pragma unroll
  for (int c = 0; c < NSTEP; ++c) {
    const int tap = 4 * c + lg < NTAP ? 4 * c + lg : NTAP - 1;
    boff[c] = ((wr0 + tap / KS) * WP + li + tap % KS) * CH;
  }
  const int cb = tile_channel<MT>(co0, 0, 4 * lg);
  float bvv[4 * MT];
  load_bias<MT>(bias, cb, bvv);
  v4f acc[MT][RW][NT];
  op a[2][MT];
  op bq[2][NT];
  for (int k = 0; k < mine; ++k) {
    __syncthreads();
    const int it = blockIdx.x + k * gridDim.x;
    const int rest = it / ncog;
    const int b = rest / nrg, row0 = (rest % nrg) * ROWS;
    const bf16_t* buf = lin + (k & 1) * BUF;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int rr = 0; rr < RW; ++rr)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[m][rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
    auto read_a = [&](op (&av)[MT], int c) {
#pragma unroll
      for (int m = 0; m < MT; ++m)
        av[m] = *reinterpret_cast<const op*>(lw + ((m * NSTEP + c) * 64 + lane) * 8);
    };
    auto read_b = [&](op (&bv)[NT], int c, int rr) {
      const bf16_t* p = buf + boff[c] + rr * WP * CH;
#pragma unroll
      for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const op*>(p + t * 16 * CH);
    };
    read_a(a[0], 0);
    read_b(bq[0], 0, 0);
    static_for<0, NSTEP * RW>([&](auto ic) {
      constexpr int idx = decltype(ic)::value;
      constexpr int c = idx / RW, rr = idx % RW;
      constexpr bool lda = rr == 0 && c + 1 < NSTEP;
      constexpr bool ldb = idx + 1 < NSTEP * RW;
      if constexpr (lda) read_a(a[(c + 1) & 1], c + 1);
      if constexpr (ldb) read_b(bq[(idx + 1) & 1], (idx + 1) / RW, (idx + 1) % RW);
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < MT; ++m)
          acc[m][rr][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[c & 1][m], bq[idx & 1][t],
                                                                  acc[m][rr][t], 0, 0, 0);
      __builtin_amdgcn_sched_group_barrier(0x100, (lda ? MT : 0) + (ldb ? NT : 0), 0);
      __builtin_amdgcn_sched_group_barrier(0x008, MT * NT, 0);
    });
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int row = row0 + wr0 + rr;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const size_t o = (((size_t)b * H + row) * W + t * 16 + li) * coutp + cb;
        float v[4 * MT];
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
          for (int j = 0; j < 4; ++j) v[4 * m + j] = acc[m][rr][t][j];
        store_channels<MT>(v, bvv, bias != nullptr, relu, nullptr, y, y32, o);
      }
    }
  }
}

// w8[co][tap][c] = w[co][tap][c] for tap < 25 and c < 8, else 0
__global__ void k_pack_first(const bf16_t* __restrict__ w, bf16_t* __restrict__ w8, int cinp,
                             int coutp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= coutp * kFirstTaps * kFirstCh) return;
  const int c = i % kFirstCh, tap = (i / kFirstCh) % kFirstTaps, co = i / (kFirstCh * kFirstTaps);
  w8[i] = (tap < NTAP && c < cinp) ? w[((size_t)co * NTAP + tap) * cinp + c] : (bf16_t)0;
}

// --------------------------------------------------------------------------
// Last layer (cout output channels with 5 * cout <= 16): the horizontal taps go
// into the M dimension.  Row c*5 + j of the MFMA tile is output channel c's
// column-offset-j part D[c, j](p) = sum over dy, ci of W[c][dy*5 + j][ci]
// x[row + dy - 2][p][ci] (K = 5 vertical taps x the input channels, no
// horizontal shift), and y[p][c] = bias[c] + sum_j D[c, j](p + j - 2), the
// horizontal sum taken through LDS.  Per pixel tile 5 * cinp/32 matrix steps
// instead of 25 * cinp/32 for a 16-channel tile of which cout are real.
// Weights repacked once per update: wl[16][5][cinp] (rows >= 5 cout zero).
// --------------------------------------------------------------------------
template <int W, int RW>
__global__ __launch_bounds__(256) void k_conv_last(const bf16_t* __restrict__ x,
                                                   const bf16_t* __restrict__ wl,
                                                   const float* __restrict__ bias, int H,
                                                   int cinp, int coutp, int cout,
                                                   bf16_t* __restrict__ y,
                                                   float* __restrict__ y32) {
  typedef v8bf op;
  constexpr int KC = 32;
  constexpr int NT = W / 16;
  constexpr int ROWS = 4 * RW;
  constexpr int SR = ROWS + KS - 1;
  constexpr int PS = KC + 8;
  constexpr int PIECES = KC / 8;
  __shared__ __attribute__((aligned(16))) bf16_t lds[SR * W * PS];
  __shared__ float part[4][RW][16][W + 4];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int row0 = blockIdx.x * ROWS, b = blockIdx.y;
  const int wrow0 = row0 + wave * RW;
  const bool wave_ok = wrow0 < H;
  const int li = lane % 16, lg = lane / 16;
  v4f acc[RW][NT];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int ci0 = 0; ci0 < cinp; ci0 += KC) {
    __syncthreads();
    stage_rows<SR, W, PIECES, PS>(x, lds, b, H, row0, cinp, ci0);
    __syncthreads();
    if (wave_ok) {
#pragma unroll
      for (int dy = 0; dy < KS; ++dy) {
        const op a = *reinterpret_cast<const op*>(wl + ((size_t)li * KS + dy) * cinp + ci0 +
                                                  lg * 8);
#pragma unroll
        for (int rr = 0; rr < RW; ++rr) {
          const bf16_t* lrow = lds + (wave * RW + rr + dy) * W * PS + lg * 8;
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const op bv = *reinterpret_cast<const op*>(lrow + (t * 16 + li) * PS);
            acc[rr][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bv, acc[rr][t], 0, 0, 0);
          }
        }
      }
    }
  }
  if (!wave_ok) return;
  // D[4 lg + j][p] of the wave's rows into LDS (its own region), then the
  // horizontal sums; every lane handles one pixel
#pragma unroll
  for (int rr = 0; rr < RW; ++rr)
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) part[wave][rr][4 * lg + j][t * 16 + li] = acc[rr][t][j];
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes are done
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int rr = 0; rr < RW; ++rr) {
    const int row = wrow0 + rr;
    if (row >= H) break;
    const int p = lane;
    const size_t o = (((size_t)b * H + row) * W + p) * coutp;
    for (int c0 = 0; c0 < coutp; c0 += 4) {
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int c = c0 + i;
        v[i] = bias ? bias[c] : 0.f;
        if (c < cout) {
          float sum = 0.f;
#pragma unroll
          for (int j = 0; j < KS; ++j) sum += part[wave][rr][c * KS + j][(p + j - 2 + W) % W];
          v[i] += sum;
        }
      }
      if (y) {
        v4s out;
#pragma unroll
        for (int i = 0; i < 4; ++i) out[i] = (short)f2bf(v[i]);
        *reinterpret_cast<v4s*>(y + o + c0) = out;
      }
      if (y32) *reinterpret_cast<v4f*>(y32 + o + c0) = v4f{v[0], v[1], v[2], v[3]};
    }
  }
}

// Last layer with overlapped staging (64-pixel rows, heights a multiple of
// 8): the work of k_conv_last, one resident block per CU walking its share of
// the (sample, 8-row group) items as cinp/32 stages.  Four loader waves stage
// the next stage's 12 input rows ([row][K group][pixel], 16-byte entries; no
// halo: the horizontal taps are summed in the epilogue) into one of two LDS
// buffers while the four compute waves run the current one (barrier pattern
// of k_conv_fwd_ws); the compute waves' stores drain under the next stage.
template <int W, int NLW>
__global__ __launch_bounds__(256 + 64 * NLW, 1) void k_conv_last_ws(
    const bf16_t* __restrict__ x, const bf16_t* __restrict__ wl, const float* __restrict__ bias,
    int B, int H, int cinp, int coutp, int cout, bf16_t* __restrict__ y,
    float* __restrict__ y32) {
  typedef v8bf op;
  constexpr int KC = 32, RW = 2, NT = W / 16;
  constexpr int ROWS = 4 * RW, SR = ROWS + KS - 1, PIECES = KC / 8;
  constexpr int BUF = SR * PIECES * W * 8;
  __shared__ __attribute__((aligned(16))) bf16_t lds[2 * BUF];
  __shared__ float part[4][RW][16][W + 4];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int nrg = H / ROWS, nch = cinp / KC;
  const int items = B * nrg;
  const int mine = (int)blockIdx.x < items ? (items - 1 - (int)blockIdx.x) / (int)gridDim.x + 1 : 0;
  const int S = mine * nch;
  if (wave >= 4) {  // loader waves: stage s into buffer s%2, then barrier s
    constexpr int NL = 64 * NLW;
    constexpr int K = PIECES * W / NL;
    static_assert(K * NL == PIECES * W, "whole rows per loader passes");
    const int lt = threadIdx.x - 256;
    for (int s = 0; s < S; ++s) {
      const int it = blockIdx.x + (s / nch) * gridDim.x;
      const int ci0 = (s % nch) * KC;
      const int b = it / nrg, row0 = (it % nrg) * ROWS;
      uint4 v[SR * K];
#pragma unroll
      for (int r = 0; r < SR; ++r) {
        const int yy = (row0 + r - 2 + H) % H;
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const int qq = lt + NL * j;
          const int pix = ((qq / 64) % (W / 16)) * 16 + qq % 16, pc = (qq / 16) % PIECES;
          v[r * K + j] = *reinterpret_cast<const uint4*>(
              x + (((size_t)b * H + yy) * W + pix) * cinp + ci0 + pc * 8);
        }
      }
      bf16_t* buf = lds + (s & 1) * BUF;
#pragma unroll
      for (int r = 0; r < SR; ++r)
#pragma unroll
        for (int j = 0; j < K; ++j) {
          const int qq = lt + NL * j;
          const int pix = ((qq / 64) % (W / 16)) * 16 + qq % 16, pc = (qq / 16) % PIECES;
          *reinterpret_cast<uint4*>(buf + ((r * PIECES + pc) * W + pix) * 8) = v[r * K + j];
        }
      __syncthreads();
    }
    return;
  }
  const int li = lane % 16, lg = lane / 16;
  const int wr0 = wave * RW;
  v4f acc[RW][NT];
#pragma unroll
  for (int rr = 0; rr < RW; ++rr)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int s = 0; s < S; ++s) {
    __syncthreads();
    const int it = blockIdx.x + (s / nch) * gridDim.x;
    const int ch = s % nch;
    const int b = it / nrg, row0 = (it % nrg) * ROWS;
    const bf16_t* buf = lds + (s & 1) * BUF;
    op a[KS];
#pragma unroll
    for (int dy = 0; dy < KS; ++dy)
      a[dy] = *reinterpret_cast<const op*>(wl + ((size_t)li * KS + dy) * cinp + ch * KC + lg * 8);
#pragma unroll
    for (int dy = 0; dy < KS; ++dy)
#pragma unroll
      for (int rr = 0; rr < RW; ++rr) {
        const bf16_t* lrow = buf + ((wr0 + rr + dy) * PIECES + lg) * W * 8 + li * 8;
        op bv[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t] = *reinterpret_cast<const op*>(lrow + t * 16 * 8);
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[rr][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[dy], bv[t], acc[rr][t], 0, 0, 0);
      }
    if (ch != nch - 1) continue;
    // D[4 lg + j][p] of the wave's rows into its own LDS region, then the
    // horizontal sums; every lane handles one pixel
#pragma unroll
    for (int rr = 0; rr < RW; ++rr)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int j = 0; j < 4; ++j) part[wave][rr][4 * lg + j][t * 16 + li] = acc[rr][t][j];
        acc[rr][t] = v4f{0.f, 0.f, 0.f, 0.f};
      }
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes are done
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int rr = 0; rr < RW; ++rr) {
      const int row = row0 + wr0 + rr;
      const int p = lane;
      const size_t o = (((size_t)b * H + row) * W + p) * coutp;
      for (int c0 = 0; c0 < coutp; c0 += 4) {
        float v[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int c = c0 + i;
          v[i] = bias ? bias[c] : 0.f;
          if (c < cout) {
            float sum = 0.f;
#pragma unroll
            for (int j = 0; j < KS; ++j) sum += part[wave][rr][c * KS + j][(p + j - 2 + W) % W];
            v[i] += sum;
          }
        }
        if (y) {
          v4s out;
#pragma unroll
          for (int i = 0; i < 4; ++i) out[i] = (short)f2bf(v[i]);
          *reinterpret_cast<v4s*>(y + o + c0) = out;
        }
        if (y32) *reinterpret_cast<v4f*>(y32 + o + c0) = v4f{v[0], v[1], v[2], v[3]};
      }
    }
    // the next item's partial sums go to the same region: every lane has
    // read this item's before any lane writes
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
  }
}

// wl[r][dy][ci] = w[r / 5][dy * 5 + r % 5][ci] for r < 5 cout, else 0
__global__ void k_pack_last(const bf16_t* __restrict__ w, bf16_t* __restrict__ wl, int cinp,
                            int cout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 16 * KS * cinp) return;
  const int ci = i % cinp, dy = (i / cinp) % KS, r = i / (cinp * KS);
  const int c = r / KS, j = r % KS;
  wl[i] = c < cout ? w[((size_t)c * NTAP + dy * KS + j) * cinp + ci] : (bf16_t)0;
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

// --------------------------------------------------------------------------
// Weight gradient, blocked for the matrix cores.  dW[co][tap][ci] = sum over
// pixels p of dy[p][co] * x[p + shift(tap)][ci]: a GEMM per tap with the pixels
// as K.  A block of WC x WI waves owns the tile co [16 WC) x ci [16 WI) for all
// 25 taps (each wave one 16 x 16 tile, 25 accumulators) and the rows
// [r_begin, r_end) of one sample.  Per output row it stages dy (transposed to
// [co][px]) and one new x row (transposed to [ci][px] with two periodic pixels
// of padding on each side) into a five-row LDS ring, the row's loads issued one
// row ahead so that they overlap the matrix ops.  For each 32-pixel K step
// one A read (dy) and two aligned 16-byte B reads per x row give the B
// operands of all five horizontal shifts (byte-aligned in registers), so one
// dy read feeds 25 MFMAs.  The block's sums go to its own slice of a workspace
// ([split][16x16 tile][tap][lane][4]) that k_wgrad_partial and k_wgrad_reduce
// add into dW.
// --------------------------------------------------------------------------
__device__ __forceinline__ v8bf window8(const uint4& c0, const uint4& c1, int dx) {
  unsigned w[4];
  const unsigned u[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  const int d = dx >> 1;
  if (dx & 1) {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = __builtin_amdgcn_alignbyte(u[d + k + 1], u[d + k], 2);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = u[d + k];
  }
  v8bf r;
  __builtin_memcpy(&r, w, sizeof(r));
  return r;
}

template <int W, int WC, int WI>
__global__ __launch_bounds__(256) void k_conv_wgrad_mt(const bf16_t* __restrict__ dy,
                                                       const bf16_t* __restrict__ x, int H,
                                                       int cinp, int coutp, int rsplit,
                                                       float* __restrict__ ws) {
  constexpr int CO = 16 * WC, CI = 16 * WI;
  constexpr int LS = W + 8;  // LDS row stride (elements): 64 + 8 keeps 16-lane reads conflict-free
  constexpr int XP = W + 8;  // padded x row: element e holds pixel (e - 2) mod W
  __shared__ __attribute__((aligned(16))) bf16_t sdy[CO * LS];
  __shared__ __attribute__((aligned(16))) bf16_t sx[5 * CI * LS];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int tiles_ci = cinp / CI;
  const int cot = blockIdx.x / tiles_ci, cit = blockIdx.x % tiles_ci;
  const int co_base = cot * CO, ci_base = cit * CI;
  const int split = blockIdx.y;
  const int b = split / rsplit, rs = split % rsplit;
  const int rows = H / rsplit;
  const int r_begin = rs * rows, r_end = r_begin + rows;
  const int wc = wave % WC, wi = wave / WC;
  const int li = lane % 16, lg = lane / 16;
  v4f acc[NTAP];
#pragma unroll
  for (int t = 0; t < NTAP; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};

  // Staging: thread tid holds x pieces tid + NTH i (16 bytes: 8 channels of
  // one padded pixel) and dy pieces tid + NTH i (8 channels of one pixel).
  // All of a row's loads are issued before any of its LDS writes, and the
  // next row's loads are in flight during this row's matrix ops.
  constexpr int NTH = 64 * WC * WI;
  constexpr int XPC = XP * (CI / 8), DPC = W * (CO / 8);
  constexpr int NXQ = (XPC + NTH - 1) / NTH, NDQ = (DPC + NTH - 1) / NTH;
  typedef unsigned v4u __attribute__((ext_vector_type(4)));
  v4u xv[NXQ], dv[NDQ];
  const int tid = threadIdx.x;
  auto load_x = [&](int q) {
    const int yy = ((q % H) + H) % H;
    static_for<0, NXQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int pc = tid + NTH * i;
      if (XPC % NTH == 0 || pc < XPC) {
        const int e = pc / (CI / 8), c8 = (pc % (CI / 8)) * 8;
        const int px = (e - 2 + W) % W;
        xv[i] = *reinterpret_cast<const v4u*>(x + (((size_t)b * H + yy) * W + px) * cinp +
                                              ci_base + c8);
      }
    });
  };
  // x row q into ring slot q mod 5
  auto write_x = [&](int q) {
    const int slot = ((q % 5) + 5) % 5;
    static_for<0, NXQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int pc = tid + NTH * i;
      if (XPC % NTH == 0 || pc < XPC) {
        const int e = pc / (CI / 8), c8 = (pc % (CI / 8)) * 8;
        const bf16_t* hv = reinterpret_cast<const bf16_t*>(&xv[i]);
#pragma unroll
        for (int c = 0; c < 8; ++c) sx[(slot * CI + c8 + c) * LS + e] = hv[c];
      }
    });
  };
  auto load_dy = [&](int r) {
    static_for<0, NDQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int pc = tid + NTH * i;
      if (DPC % NTH == 0 || pc < DPC) {
        const int px = pc / (CO / 8), c8 = (pc % (CO / 8)) * 8;
        dv[i] = *reinterpret_cast<const v4u*>(dy + (((size_t)b * H + r) * W + px) * coutp +
                                              co_base + c8);
      }
    });
  };
  auto write_dy = [&]() {
    static_for<0, NDQ>([&](auto ic) {
      constexpr int i = decltype(ic)::value;
      const int pc = tid + NTH * i;
      if (DPC % NTH == 0 || pc < DPC) {
        const int px = pc / (CO / 8), c8 = (pc % (CO / 8)) * 8;
        const bf16_t* hv = reinterpret_cast<const bf16_t*>(&dv[i]);
#pragma unroll
        for (int c = 0; c < 8; ++c) sdy[(c8 + c) * LS + px] = hv[c];
      }
    });
  };
  for (int q = r_begin - 2; q < r_begin + 2; ++q) {
    load_x(q);
    write_x(q);
  }
  load_x(r_begin + 2);
  load_dy(r_begin);
  for (int r = r_begin; r < r_end; ++r) {
    __syncthreads();  // previous row's reads are done
    write_x(r + 2);
    write_dy();
    __syncthreads();
    if (r + 1 < r_end) {
      load_x(r + 3);
      load_dy(r + 1);
    }
#pragma unroll
    for (int kk = 0; kk < W / 32; ++kk) {
      const int p0 = kk * 32 + lg * 8;
      const v8bf a = *reinterpret_cast<const v8bf*>(sdy + (wc * 16 + li) * LS + p0);
#pragma unroll
      for (int t = 0; t < KS; ++t) {
        const int slot = (((r + t - 2) % 5) + 5) % 5;
        const bf16_t* xr = sx + (slot * CI + wi * 16 + li) * LS + p0;
        const uint4 c0 = *reinterpret_cast<const uint4*>(xr);
        const uint4 c1 = *reinterpret_cast<const uint4*>(xr + 8);
#pragma unroll
        for (int dx = 0; dx < KS; ++dx)
          acc[t * KS + dx] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, window8(c0, c1, dx),
                                                                     acc[t * KS + dx], 0, 0, 0);
      }
    }
  }
  // lane holds D[co = 4 lg + j][ci = li] of its wave's tile, for every tap:
  // the workspace keeps it as [16x16 tile][tap][lane][j], so that each lane
  // stores 16 bytes per tap and the wave 1 KB in a row
  const int tile = (cot * WC + wc) * (cinp / 16) + cit * WI + wi;
  float* out = ws + (size_t)split * NTAP * coutp * cinp + (size_t)tile * NTAP * 256;
#pragma unroll
  for (int t = 0; t < NTAP; ++t) *reinterpret_cast<v4f*>(out + (t * 64 + lane) * 4) = acc[t];
}

// First stage of the workspace reduction: part[g][i] = sum of ws[k][i] over
// the splits k of group g (blockIdx.y), in a fixed order.
constexpr int kSplitGroup = 32;
__global__ void k_wgrad_partial(const float* __restrict__ ws, float* __restrict__ part,
                                int nsplit, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k0 = blockIdx.y * kSplitGroup;
  const int k1 = k0 + kSplitGroup < nsplit ? k0 + kSplitGroup : nsplit;
  float s = 0.f;
  int k = k0;
  // eight loads in flight, added in the same order as one at a time
  for (; k + 8 <= k1; k += 8) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = ws[(size_t)(k + j) * n + i];
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  for (; k < k1; ++k) s += ws[(size_t)k * n + i];
  part[(size_t)blockIdx.y * n + i] = s;
}

// dW[co][tap][ci] += sum over groups of part[g][tap][co][ci]
__global__ void k_wgrad_reduce(const float* __restrict__ part, float* __restrict__ dw,
                               int ngroups, int cinp, int coutp) {
  const size_t n = (size_t)NTAP * coutp * cinp;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    // workspace order [16x16 (co, ci) tile][tap][lane][j]: lane = li + 16 lg
    // holds co = 16 tco + 4 lg + j, ci = 16 tci + li
    const int j = (int)(i % 4), lane = (int)((i / 4) % 64);
    const int tap = (int)((i / 256) % NTAP), tile = (int)(i / (256 * NTAP));
    const int co = (tile / (cinp / 16)) * 16 + (lane / 16) * 4 + j;
    const int ci = (tile % (cinp / 16)) * 16 + lane % 16;
    float s = 0.f;
    for (int g = 0; g < ngroups; ++g) s += part[(size_t)g * n + i];
    dw[((size_t)co * NTAP + tap) * cinp + ci] += s;
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

// db[c] += sum over pixels of dy[px][c]: each thread sums one channel over a
// strided set of pixels, the block combines its threads in LDS, and one atomic
// per channel and block adds the result (blockDim = 256, a multiple of coutp).
__global__ __launch_bounds__(256) void k_bias_grad(const bf16_t* __restrict__ dy, size_t npix,
                                                   int coutp, float* __restrict__ db) {
  __shared__ float red[256];
  const int c = threadIdx.x % coutp;
  const int lanes_per_c = blockDim.x / coutp;
  const int sub = threadIdx.x / coutp;
  float s = 0.f;
  const size_t stride = (size_t)gridDim.x * lanes_per_c;
  size_t p = blockIdx.x * (size_t)lanes_per_c + sub;
  // eight loads in flight, added in the same order as one at a time
  for (; p + 7 * stride < npix; p += 8 * stride) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = bf2f(dy[(p + j * stride) * coutp + c]);
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  for (; p < npix; p += stride) s += bf2f(dy[p * coutp + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (sub == 0) {
    for (int k = 1; k < lanes_per_c; ++k) s += red[k * coutp + c];
    atomicAdd(db + c, s);
  }
}

// The same sum in a fixed order: each block writes its per-channel partial
// sums to part[block][c] (no atomics), then k_bias_grad_sum adds the blocks'
// partials channel by channel in block order, so repeated runs agree bit for bit.
__global__ __launch_bounds__(256) void k_bias_grad_part(const bf16_t* __restrict__ dy,
                                                        size_t npix, int coutp,
                                                        float* __restrict__ part) {
  __shared__ float red[256];
  const int c = threadIdx.x % coutp;
  const int lanes_per_c = blockDim.x / coutp;
  const int sub = threadIdx.x / coutp;
  float s = 0.f;
  const size_t stride = (size_t)gridDim.x * lanes_per_c;
  size_t p = blockIdx.x * (size_t)lanes_per_c + sub;
  // eight loads in flight, added in the same order as one at a time
  for (; p + 7 * stride < npix; p += 8 * stride) {
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = bf2f(dy[(p + j * stride) * coutp + c]);
#pragma unroll
    for (int j = 0; j < 8; ++j) s += v[j];
  }
  for (; p < npix; p += stride) s += bf2f(dy[p * coutp + c]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (sub == 0) {
    for (int k = 1; k < lanes_per_c; ++k) s += red[k * coutp + c];
    part[(size_t)blockIdx.x * coutp + c] = s;
  }
}

// one block per channel: the nblocks (<= 256) partials loaded side by side and
// added by a tree in LDS, whose order is fixed
__global__ __launch_bounds__(256) void k_bias_grad_sum(const float* __restrict__ part,
                                                       int nblocks, int coutp,
                                                       float* __restrict__ db) {
  __shared__ float red[256];
  const int c = blockIdx.x;
  red[threadIdx.x] = threadIdx.x < nblocks ? part[(size_t)threadIdx.x * coutp + c] : 0.f;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] += red[0];
}

// MSE over the first `nch` of `coutp` channels: loss += sum (pred - target)^2 * inv_n;
// grad (bf16, all coutp channels, zero in the padding) = 2 (pred - target) * inv_n * gscale
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      size_t npix, int coutp, int nch, float inv_n, float gscale,
                      bf16_t* __restrict__ grad, float* __restrict__ loss) {
  float s = 0.f;
  if (coutp % 8 == 0) {
    // 8 channels of one pixel per thread and step: 32-byte reads and a
    // 16-byte gradient store, consecutive lanes on consecutive pieces
    const int cg = coutp / 8;
    for (size_t i8 = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i8 < npix * cg;
         i8 += (size_t)gridDim.x * blockDim.x) {
      const size_t p = i8 / cg;
      const int c0 = (int)(i8 - p * cg) * 8;
      const v4f a0 = *reinterpret_cast<const v4f*>(pred + i8 * 8);
      const v4f a1 = *reinterpret_cast<const v4f*>(pred + i8 * 8 + 4);
      v8s g8;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        float g = 0.f;
        if (c0 + j < nch) {
          const float d = (j < 4 ? a0[j] : a1[j - 4]) - target[p * nch + c0 + j];
          s += d * d;
          g = 2.f * d * inv_n * gscale;
        }
        g8[j] = (short)f2bf(g);
      }
      if (grad) *reinterpret_cast<v8s*>(grad + i8 * 8) = g8;
    }
  } else {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < npix;
         p += (size_t)gridDim.x * blockDim.x) {
      for (int c = 0; c < coutp; ++c) {
        float g = 0.f;
        if (c < nch) {
          const float d = pred[p * coutp + c] - target[p * nch + c];
          s += d * d;
          g = 2.f * d * inv_n * gscale;
        }
        if (grad) grad[p * coutp + c] = f2bf(g);
      }
    }
  }
  // wave reduction, then the block's waves through LDS, one atomic per block
  __shared__ float red[16];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0 && loss) {
    float t = 0.f;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    atomicAdd(loss, t * inv_n);
  }
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

// The convolution kernels read and write 16-byte vectors (inputs, weights,
// bias, mask, outputs) at offsets that are multiples of 16 bytes from each base.
bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }
bool all_aligned16(std::initializer_list<const void*> ps) {
  for (const void* p : ps)
    if (!aligned16(p)) return false;
  return true;
}

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
  if (!all_aligned16({x, w, bias, mask, y, y32}))
    return fail(QGCNN_E_ARG, "tensors must start on 16-byte boundaries");
  hipStream_t st = (hipStream_t)stream;
  const int mt = coutp % 64 == 0 ? 4 : (coutp % 32 == 0 ? 2 : 1);
  if (cinp % 32 == 0 && H % 8 == 0 && mt == 4 &&
      (size_t)B * H * W * (cinp > coutp ? cinp : coutp) * sizeof(bf16_t) < (1u << 31)) {
    // 64-channel output groups: overlapped staging, one resident block per CU
    // over the layer's work items (128 -> 64 channels: 189 against 235 us)
    const int items = B * (H / 8) * (coutp / 64);
    const int grid = items < num_cus() ? items : num_cus();
    hipLaunchKernelGGL((k_conv_fwd_ws<64, 4, 2, kLoaderWaves, true>), dim3(grid),
                       dim3(256 + 64 * kLoaderWaves), 0, st, (const bf16_t*)x, (const bf16_t*)w,
                       bias, B, H, cinp, coutp, relu, (const bf16_t*)mask, (bf16_t*)y, y32);
    HIPC(hipGetLastError());
    return 0;
  }
  if (cinp % 32 == 0 && H % 8 == 0 && mt == 2 &&
      (size_t)B * H * W * (cinp > coutp ? cinp : coutp) * sizeof(bf16_t) < (1u << 31)) {
    // 32-channel output groups: overlapped staging with two resident blocks
    // per CU, each with one LDS stage (51 KB; 4 KB of unused dynamic LDS keep
    // a third block off the CU, so that the grid of two blocks per CU spreads
    // evenly).  32 -> 32 channels 30 against 33 us, 64 -> 32 74 against 81.
    const int items = B * (H / 8) * (coutp / 32);
    const int grid = items < 2 * num_cus() ? items : 2 * num_cus();
    hipLaunchKernelGGL((k_conv_fwd_ws<64, 2, 2, 4, false>), dim3(grid), dim3(256 + 64 * 4), 4096,
                       st, (const bf16_t*)x, (const bf16_t*)w, bias, B, H, cinp, coutp, relu,
                       (const bf16_t*)mask, (bf16_t*)y, y32);
    HIPC(hipGetLastError());
    return 0;
  }
  // two output rows per wave wherever the height allows: every weight operand
  // then feeds twice the MFMAs (measured 12 % faster at 32 -> 32 channels too,
  // although the larger LDS tile halves the blocks per CU)
  const int rw = H % 8 == 0 ? 2 : 1;
  dim3 grid((H + 4 * rw - 1) / (4 * rw), B, coutp / (16 * mt));
#define QGCNN_FWD(MT, KC, RW)                                                                  \
  hipLaunchKernelGGL((k_conv_fwd_mt<64, MT, KC, RW>), grid, dim3(256), 0, st, (const bf16_t*)x, \
                     (const bf16_t*)w, bias, H, cinp, coutp, relu, (const bf16_t*)mask,        \
                     (bf16_t*)y, y32)
#define QGCNN_FWD_RW(MT, KC) \
  do { if (rw == 2) QGCNN_FWD(MT, KC, 2); else QGCNN_FWD(MT, KC, 1); } while (0)
  if (cinp % 32 == 0) {
    if (mt == 4) QGCNN_FWD_RW(4, 32); else if (mt == 2) QGCNN_FWD_RW(2, 32); else QGCNN_FWD_RW(1, 32);
  } else {
    if (mt == 4) QGCNN_FWD_RW(4, 16); else if (mt == 2) QGCNN_FWD_RW(2, 16); else QGCNN_FWD_RW(1, 16);
  }
#undef QGCNN_FWD_RW
#undef QGCNN_FWD
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_pack_first(const void* w, int cinp, int coutp, void* w8, void* stream) {
  if (!w || !w8) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || !mult16(coutp))
    return fail(QGCNN_E_ARG, "channel counts must be multiples of 16 in [16, 128]");
  const int n = coutp * kFirstTaps * kFirstCh;
  hipLaunchKernelGGL(k_pack_first, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     (const bf16_t*)w, (bf16_t*)w8, cinp, coutp);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_conv_first(const void* x, int B, int H, int W, int cinp, const void* w8,
                     const float* bias, int coutp, int relu, void* y, float* y32, void* stream) {
  if (!x || !w8 || (!y && !y32)) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || !mult16(coutp))
    return fail(QGCNN_E_ARG, "channel counts must be multiples of 16 in [16, 128]");
  if (B < 1 || H < KS || W != 64)
    return fail(QGCNN_E_ARG, "images must be W=64 wide and at least 5 rows (got %dx%d)", H, W);
  if (!all_aligned16({x, w8, bias, y, y32}))
    return fail(QGCNN_E_ARG, "tensors must start on 16-byte boundaries");
  hipStream_t st = (hipStream_t)stream;
  const int mt = coutp % 64 == 0 ? 4 : (coutp % 32 == 0 ? 2 : 1);
  if (mt == 4 && H % 8 == 0) {
    // overlapped staging, one resident block per CU; the grid a multiple of
    // the output-channel groups so that each block keeps one group's weights
    const int ncog = coutp / 64;
    const int items = B * (H / 8) * ncog;
    int grid = items < num_cus() ? items : num_cus() / ncog * ncog;
    if (grid < ncog) grid = ncog;
    hipLaunchKernelGGL((k_conv_first_ws<64, kLoaderWaves>), dim3(grid),
                       dim3(256 + 64 * kLoaderWaves), 0, st, (const bf16_t*)x, (const bf16_t*)w8,
                       bias, B, H, cinp, coutp, relu, (bf16_t*)y, y32);
    HIPC(hipGetLastError());
    return 0;
  }
  const int rw = H % 8 == 0 ? 2 : 1;
  dim3 grid((H + 4 * rw - 1) / (4 * rw), B, coutp / (16 * mt));
#define QGCNN_FIRST(MT, RW)                                                                    \
  hipLaunchKernelGGL((k_conv_first<64, MT, RW>), grid, dim3(256), 0, st, (const bf16_t*)x,    \
                     (const bf16_t*)w8, bias, H, cinp, coutp, relu, (bf16_t*)y, y32)
  if (rw == 2) {
    if (mt == 4) QGCNN_FIRST(4, 2); else if (mt == 2) QGCNN_FIRST(2, 2); else QGCNN_FIRST(1, 2);
  } else {
    if (mt == 4) QGCNN_FIRST(4, 1); else if (mt == 2) QGCNN_FIRST(2, 1); else QGCNN_FIRST(1, 1);
  }
#undef QGCNN_FIRST
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_pack_last(const void* w, int cinp, int coutp, int cout, void* wl, void* stream) {
  if (!w || !wl) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || cinp % 32 != 0 || !mult16(coutp) || cout < 1 || KS * cout > 16 ||
      cout > coutp)
    return fail(QGCNN_E_ARG, "last-layer packing needs cinp a multiple of 32 and 5*cout <= 16");
  const int n = 16 * KS * cinp;
  hipLaunchKernelGGL(k_pack_last, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     (const bf16_t*)w, (bf16_t*)wl, cinp, cout);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_conv_last(const void* x, int B, int H, int W, int cinp, const void* wl,
                    const float* bias, int coutp, int cout, void* y, float* y32, void* stream) {
  if (!x || !wl || (!y && !y32)) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || cinp % 32 != 0 || coutp % 4 != 0 || coutp < 4 || coutp > 128 ||
      cout < 1 || KS * cout > 16 || cout > coutp)
    return fail(QGCNN_E_ARG, "the last-layer kernel needs cinp a multiple of 32, 5*cout <= 16 "
                             "and coutp a multiple of 4 in [4, 128]");
  if (!all_aligned16({x, wl, bias, y32}) || ((uintptr_t)y & 7u))
    return fail(QGCNN_E_ARG, "tensors must start on 16-byte boundaries (y on 8-byte ones)");
  if (B < 1 || H < KS || W != 64)
    return fail(QGCNN_E_ARG, "images must be W=64 wide and at least 5 rows (got %dx%d)", H, W);
  if (H % 8 == 0) {
    // overlapped staging, one resident block per CU over the (sample, 8-row
    // group) items
    const int items = B * (H / 8);
    const int grid = items < num_cus() ? items : num_cus();
    hipLaunchKernelGGL((k_conv_last_ws<64, kLoaderWaves>), dim3(grid),
                       dim3(256 + 64 * kLoaderWaves), 0, (hipStream_t)stream, (const bf16_t*)x,
                       (const bf16_t*)wl, bias, B, H, cinp, coutp, cout, (bf16_t*)y, y32);
    HIPC(hipGetLastError());
    return 0;
  }
  dim3 grid((H + 3) / 4, B);
  hipLaunchKernelGGL((k_conv_last<64, 1>), grid, dim3(256), 0, (hipStream_t)stream,
                     (const bf16_t*)x, (const bf16_t*)wl, bias, H, cinp, coutp, cout, (bf16_t*)y,
                     y32);
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

namespace {

// wave layout and K split of the blocked weight-gradient kernel
struct WgradPlan {
  int wc, wi, rsplit, nsplit, tiles;
};

WgradPlan wgrad_plan(int B, int H, int cinp, int coutp) {
  WgradPlan p;
  // waves per block along co and ci: counts that divide the 16-channel tiles
  const int nco = coutp / 16, nci = cinp / 16;
  p.wc = nco % 2 == 0 ? 2 : 1;
  p.wi = nci % 2 == 0 ? 2 : 1;
  if (p.wi == 1 && nco % 4 == 0) p.wc = 4;
  if (p.wc == 1 && nci % 4 == 0) p.wi = 4;
  p.tiles = (coutp / (16 * p.wc)) * (cinp / (16 * p.wi));
  // rows split until about two blocks per CU: each block writes a partial
  // of 25 x coutp x cinp floats, which the reductions read back, so more
  // splits cost more than they hide
  p.rsplit = 1;
  while (p.rsplit < 8 && H % (2 * p.rsplit) == 0 && H / (2 * p.rsplit) >= 4 &&
         (size_t)p.tiles * B * p.rsplit < 512)
    p.rsplit *= 2;
  p.nsplit = B * p.rsplit;
  return p;
}

}  // namespace

size_t qgcnn_wgrad_workspace_bytes(int B, int H, int W, int cinp, int coutp) {
  if (!mult16(cinp) || !mult16(coutp) || B < 1 || H < KS || W != 64) return 0;
  const WgradPlan p = wgrad_plan(B, H, cinp, coutp);
  const size_t groups = (p.nsplit + kSplitGroup - 1) / kSplitGroup;
  return ((size_t)p.nsplit + groups) * NTAP * coutp * cinp * sizeof(float);
}

int qgcnn_conv_wgrad_ws(const void* dy, const void* x, int B, int H, int W, int cinp, int coutp,
                        float* dw, void* workspace, size_t ws_bytes, void* stream) {
  if (!dy || !x || !dw || !workspace) return fail(QGCNN_E_ARG, "NULL argument");
  if (!mult16(cinp) || !mult16(coutp))
    return fail(QGCNN_E_ARG, "channel counts must be multiples of 16 in [16, 128]");
  if (B < 1 || H < KS || W != 64) return fail(QGCNN_E_ARG, "images must be W=64 wide");
  const WgradPlan p = wgrad_plan(B, H, cinp, coutp);
  const size_t need = qgcnn_wgrad_workspace_bytes(B, H, W, cinp, coutp);
  if (ws_bytes < need)
    return fail(QGCNN_E_ARG, "workspace of %zu bytes is smaller than %zu", ws_bytes, need);
  if (H % p.rsplit != 0) return fail(QGCNN_E_ARG, "internal row split does not divide H");
  hipStream_t st = (hipStream_t)stream;
  dim3 grid(p.tiles, p.nsplit);
  dim3 block(64 * p.wc * p.wi);
#define QGCNN_WG(WC, WI)                                                                  \
  hipLaunchKernelGGL((k_conv_wgrad_mt<64, WC, WI>), grid, block, 0, st, (const bf16_t*)dy, \
                     (const bf16_t*)x, H, cinp, coutp, p.rsplit, (float*)workspace)
  if (p.wc == 2 && p.wi == 2) QGCNN_WG(2, 2);
  else if (p.wc == 4 && p.wi == 1) QGCNN_WG(4, 1);
  else if (p.wc == 1 && p.wi == 4) QGCNN_WG(1, 4);
  else if (p.wc == 2 && p.wi == 1) QGCNN_WG(2, 1);
  else if (p.wc == 1 && p.wi == 2) QGCNN_WG(1, 2);
  else QGCNN_WG(1, 1);
#undef QGCNN_WG
  HIPC(hipGetLastError());
  const size_t n = (size_t)NTAP * coutp * cinp;
  const int ngroups = (p.nsplit + kSplitGroup - 1) / kSplitGroup;
  float* part = (float*)workspace + (size_t)p.nsplit * n;
  hipLaunchKernelGGL(k_wgrad_partial, dim3((unsigned)((n + 255) / 256), ngroups), dim3(256), 0,
                     st, (const float*)workspace, part, p.nsplit, n);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3(grid1(n)), dim3(256), 0, st, (const float*)part, dw,
                     ngroups, cinp, coutp);
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
  hipLaunchKernelGGL(k_bias_grad, dim3(256), dim3(256 / coutp * coutp), 0, (hipStream_t)stream,
                     (const bf16_t*)dy, npix, coutp, db);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_bias_grad_ws(const void* dy, size_t npix, int coutp, float* db, float* ws,
                       size_t ws_floats, void* stream) {
  constexpr int kBlocks = 256;
  if (!dy || !db || !ws || !mult16(coutp)) return fail(QGCNN_E_ARG, "bad arguments");
  if (ws_floats < (size_t)kBlocks * coutp)
    return fail(QGCNN_E_ARG, "workspace of %zu floats, %d needed", ws_floats, kBlocks * coutp);
  hipLaunchKernelGGL(k_bias_grad_part, dim3(kBlocks), dim3(256 / coutp * coutp), 0,
                     (hipStream_t)stream, (const bf16_t*)dy, npix, coutp, ws);
  hipLaunchKernelGGL(k_bias_grad_sum, dim3(coutp), dim3(256), 0, (hipStream_t)stream, ws,
                     kBlocks, coutp, db);
  HIPC(hipGetLastError());
  return 0;
}

int qgcnn_mse(const float* pred, const float* target, size_t npix, int coutp, int nch,
              float gscale, void* grad, float* loss, void* stream) {
  if (!pred || !target || nch < 1 || nch > coutp) return fail(QGCNN_E_ARG, "bad arguments");
  const float inv_n = 1.0f / (float)((double)npix * nch);
  if (coutp % 8 == 0 && !all_aligned16({pred, grad}))
    return fail(QGCNN_E_ARG, "pred and grad must start on 16-byte boundaries");
  const size_t work = coutp % 8 == 0 ? npix * (coutp / 8) : npix;
  const int blocks = grid1(work) < 1024 ? grid1(work) : 1024;
  hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, (hipStream_t)stream, pred, target, npix,
                     coutp, nch, inv_n, gscale, (bf16_t*)grad, loss);
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
