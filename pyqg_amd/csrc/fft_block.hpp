// Block-level FP64 complex FFT for CDNA4 (gfx950).
//
// One transform of length N is carried by T = N/E threads; thread t holds E
// complex values in registers, x[s] = element (t + s*T).  The transform is a
// sequence of Stockham auto-sort passes of radix R | E (largest radix first),
// each pass a register-resident radix-R DFT; consecutive passes exchange data
// through LDS.  Several independent transforms ("lanes") share one LDS buffer:
// the columns of a tile interleaved with stride LANES, so a wave touches
// consecutive 16-byte words; the rows of a row group one swizzled region each.
// On exit x[s] holds output element (t + s*T) in natural order, so a forward
// transform can feed an inverse one (or the reverse) without any reordering.
// N is 2^a or 3 * 2^a; for the latter E is 6 or 12 and the first pass is the
// radix-3 one (radix 6 or 12).
//
// Twiddles come from a per-length table tw[m] = exp(-2*pi*i*m/N), built on the
// host in extended precision (correctly rounded to FP64).
#pragma once
#include <hip/hip_runtime.h>

namespace qgk {

struct cd {
  double x, y;
};

__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cmul(cd a, cd b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cd cscale(cd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cd cconj(cd a) { return {a.x, -a.y}; }

// exp(SIGN * 2*pi*i * M / R) * v for compile-time M, R.
template <int R, int M, int SIGN>
__device__ __forceinline__ cd twiddle_const(cd v) {
  constexpr int m = M % R;
  if constexpr (m == 0) {
    return v;
  } else if constexpr (4 * m == R) {  // SIGN * i
    return SIGN < 0 ? cd{v.y, -v.x} : cd{-v.y, v.x};
  } else if constexpr (2 * m == R) {
    return {-v.x, -v.y};
  } else if constexpr (4 * m == 3 * R) {  // -SIGN * i
    return SIGN < 0 ? cd{-v.y, v.x} : cd{v.y, -v.x};
  } else if constexpr ((12 * m) % R == 0) {  // multiples of 30 degrees (R = 3, 6, 12, 24)
    constexpr int k = 12 * m / R;            // 1..11, not 3, 6, 9 (handled above)
    constexpr double h = 0.86602540378443864676;  // cos(30 deg) = sqrt(3)/2
    constexpr double C12[12] = {1.0, h, 0.5, 0.0, -0.5, -h, -1.0, -h, -0.5, 0.0, 0.5, h};
    constexpr double S12[12] = {0.0, 0.5, h, 1.0, h, 0.5, 0.0, -0.5, -h, -1.0, -h, -0.5};
    constexpr double c = C12[k];
    constexpr double s = SIGN * S12[k];
    return {v.x * c - v.y * s, v.x * s + v.y * c};
  } else if constexpr (8 * m == R || 8 * m == 3 * R || 8 * m == 5 * R || 8 * m == 7 * R) {
    constexpr double h = 0.70710678118654752440;
    constexpr double c = (8 * m == R || 8 * m == 7 * R) ? h : -h;
    constexpr double s0 = (8 * m == R || 8 * m == 3 * R) ? h : -h;  // sin(2 pi m / R)
    constexpr double s = SIGN * s0;
    return {v.x * c - v.y * s, v.x * s + v.y * c};
  } else {
    // 16th roots of unity (the largest register radix); 32 only for R = 32.
    constexpr double C16[4] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                               0.38268343236508977173};
    static_assert(R == 16 || R == 32, "unsupported radix");
    if constexpr (R == 16) {
      // m in {1,3,5,7,9,11,13,15}: cos/sin from the first-quadrant table
      constexpr int q = m % 4;  // 1 or 3
      constexpr int quad = m / 4;
      constexpr double cq = q == 1 ? C16[1] : C16[3];
      constexpr double sq = q == 1 ? C16[3] : C16[1];
      constexpr double c = quad == 0 ? cq : quad == 1 ? -sq : quad == 2 ? -cq : sq;
      constexpr double s0 = quad == 0 ? sq : quad == 1 ? cq : quad == 2 ? -sq : -cq;
      constexpr double s = SIGN * s0;
      return {v.x * c - v.y * s, v.x * s + v.y * c};
    } else {
      return v;  // not reached
    }
  }
}

// In-register radix-R DFT in natural order: x[m] <- sum_s x[s] exp(SIGN 2 pi i s m / R).
template <int R, int SIGN>
struct DFT {
  template <int M>
  __device__ __forceinline__ static void combine(cd* x, const cd* e, const cd* o) {
    if constexpr (M < R / 2) {
      cd t = twiddle_const<R, M, SIGN>(o[M]);
      x[M] = cadd(e[M], t);
      x[M + R / 2] = csub(e[M], t);
      combine<M + 1>(x, e, o);
    }
  }
  __device__ __forceinline__ static void run(cd* x) {
    constexpr int H = R / 2;
    cd e[H], o[H];
#pragma unroll
    for (int i = 0; i < H; ++i) {
      e[i] = x[2 * i];
      o[i] = x[2 * i + 1];
    }
    DFT<H, SIGN>::run(e);
    DFT<H, SIGN>::run(o);
    combine<0>(x, e, o);
  }
};
template <int SIGN>
struct DFT<1, SIGN> {
  __device__ __forceinline__ static void run(cd*) {}
};
template <int SIGN>
struct DFT<2, SIGN> {
  __device__ __forceinline__ static void run(cd* x) {
    cd a = x[0];
    x[0] = cadd(a, x[1]);
    x[1] = csub(a, x[1]);
  }
};
// Radix 3: x1' = x0 + w x1 + conj(w) x2, x2' = x0 + conj(w) x1 + w x2 with
// w = exp(SIGN 2 pi i / 3) = -1/2 + SIGN i sqrt(3)/2.  Radix 6, 12 and 24 are
// the even/odd recursion above over this one.
template <int SIGN>
struct DFT<3, SIGN> {
  __device__ __forceinline__ static void run(cd* x) {
    constexpr double s = SIGN * 0.86602540378443864676;
    const cd a = cadd(x[1], x[2]);
    const cd d = csub(x[1], x[2]);
    const cd m = {x[0].x - 0.5 * a.x, x[0].y - 0.5 * a.y};
    const cd r = {-s * d.y, s * d.x};  // i s (x1 - x2)
    x[0] = cadd(x[0], a);
    x[1] = cadd(m, r);
    x[2] = csub(m, r);
  }
};
template <int SIGN>
struct DFT<4, SIGN> {
  __device__ __forceinline__ static void run(cd* x) {
    cd a0 = cadd(x[0], x[2]), a1 = csub(x[0], x[2]);
    cd a2 = cadd(x[1], x[3]);
    cd a3 = twiddle_const<4, 1, SIGN>(csub(x[1], x[3]));
    x[0] = cadd(a0, a2);
    x[2] = csub(a0, a2);
    x[1] = cadd(a1, a3);
    x[3] = csub(a1, a3);
  }
};

// LDS placement of element e of transform `lane`.
//
// Interleaved: transforms interleaved word by word (a column tile: the lanes of
// a wave are consecutive columns, so every access is a run of whole lines).
//
// at(base, off, lane) places element base + off, where `off` is a constant once
// the caller's loops are unrolled: layouts use it to fold the offset into the
// instruction's immediate.
template <int LANES>
struct Interleaved {
  __device__ __forceinline__ static int at(int e, int lane) { return e * LANES + lane; }
  __device__ __forceinline__ static int at(int base, int off, int lane) {
    return (base + off) * LANES + lane;
  }
};
// RowSwizzled: one contiguous region of N words per transform (a row); inside
// each block of 8 words the low three index bits are XORed with the block
// number, so the power-of-two lane strides of the Stockham writes, which would
// land on one bank group, spread over all of them, while runs of consecutive
// elements stay on distinct banks.
template <int N>
struct RowSwizzled {
  __device__ __forceinline__ static int swz(int e) { return e ^ ((e >> 3) & 7); }
  __device__ __forceinline__ static int at(int e, int lane) { return lane * N + swz(e); }
  // an offset that is a multiple of 64 leaves index bits 0-5 alone, so it
  // passes through the swizzle unchanged
  __device__ __forceinline__ static int at(int base, int off, int lane) {
    return (off & 63) == 0 ? lane * N + swz(base) + off : lane * N + swz(base + off);
  }
};

// One Stockham pass of radix R on sub-transforms of length L (L*R <= N).
// Butterfly b (0 <= b < N/R) reads positions b + s*N/R, writes k*L*R + j + m*L
// with j = b % L, k = b / L, after twiddling input s by w^(j*s*N/(L*R)).
template <int N, int E, int R, int L, int SIGN, class LY, bool FIRST, bool LAST, int TAG>
__device__ __forceinline__ void stockham_pass(cd (&x)[E], cd* lds, int t, int lane,
                                              const cd* __restrict__ tw) {
  constexpr int T = N / E;
  constexpr int NB = E / R;  // butterflies per thread
  cd y[NB][R];
  if constexpr (FIRST) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int s = 0; s < R; ++s) y[i][s] = x[i + s * NB];
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
#pragma unroll
      for (int s = 0; s < R; ++s) y[i][s] = lds[LY::at(t, i * T + s * (N / R), lane)];
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    if constexpr (L > 1) {
      // input s is twiddled by w^s, w = exp(SIGN 2 pi i j / (L R)): one table
      // load, the higher powers by multiplication (error grows ~ s ulp)
      const int b = t + i * T;
      const int j = b % L;
      constexpr int STEP = N / (L * R);
      cd w = tw[j * STEP];
      if constexpr (SIGN > 0) w.y = -w.y;
      cd ws = w;
#pragma unroll
      for (int s = 1; s < R; ++s) {
        y[i][s] = cmul(y[i][s], ws);
        if (s + 1 < R) ws = cmul(ws, w);
      }
    }
    DFT<R, SIGN>::run(y[i]);
  }
  if constexpr (LAST) {
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int m = 0; m < R; ++m) x[i + m * NB] = y[i][m];
  } else {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int b = t + i * T;
      const int j = b % L;
      const int k = b / L;
#pragma unroll
      for (int m = 0; m < R; ++m) lds[LY::at(k * L * R + j + m * L, lane)] = y[i][m];
    }
  }
}

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

template <int N, int E, int SIGN, class LY, int L, int TAG>
struct StockhamFrom {
  static constexpr int REM = N / L;
  // the largest radix that divides both what is left of N and E (E itself
  // while it can; N = 3 * 2^a needs E divisible by 3)
  static constexpr int R = gcd_c(REM, E);
  static_assert(R > 1, "E must hold every prime factor of N");
  __device__ __forceinline__ static void run(cd (&x)[E], cd* lds, int t, int lane,
                                             const cd* __restrict__ tw) {
    stockham_pass<N, E, R, L, SIGN, LY, (L == 1), (L * R == N), TAG>(x, lds, t, lane, tw);
    if constexpr (L * R < N) StockhamFrom<N, E, SIGN, LY, L * R, TAG>::run(x, lds, t, lane, tw);
  }
};

// Full transform.  All threads of the block must call it (it contains barriers
// whenever N > E).  lds must hold N*LANES complex values.
//
// TAG only makes distinct instantiations: a hot kernel passes its own TAG so
// that no other kernel shares its transform code.  When two kernels share one
// instantiation the compiler has generated different, worse code for one of
// them (generic instead of global/LDS addressing, and register spills).
template <int N, int E, int SIGN, int LANES, int TAG = 0>
__device__ __forceinline__ void fft_block(cd (&x)[E], cd* lds, int t, int lane,
                                          const cd* __restrict__ tw) {
  static_assert(N % E == 0, "E must divide N");
  StockhamFrom<N, E, SIGN, Interleaved<LANES>, 1, TAG>::run(x, lds, t, lane, tw);
}

// As fft_block, for transforms along rows: transform `row` of the block keeps
// its own swizzled LDS region (RowSwizzled<N>); lds must hold N*rows values.
template <int N, int E, int SIGN, int TAG = 0>
__device__ __forceinline__ void fft_block_rows(cd (&x)[E], cd* lds, int t, int row,
                                               const cd* __restrict__ tw) {
  static_assert(N % E == 0 && N % 8 == 0, "E must divide N; rows are multiples of 8");
  StockhamFrom<N, E, SIGN, RowSwizzled<N>, 1, TAG>::run(x, lds, t, row, tw);
}

}  // namespace qgk
