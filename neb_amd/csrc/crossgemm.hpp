// Split-R cross GEMM  G = sum_segments sign_s * X_s^T diag(w) Y_s  (FP64 out).
//
// This is the quadrature-weighted Gram / stiffness assembly Phi^T W Phi
// (VCORE:40-46, experiments/metaball/run.py:99-116, experiments/periodic/
// paper_rollout.py:178-179), the packed weight-gradient contraction A'^T P
// (VJ2:29, FAST:27-29, HB2P:1087-1144) and the projection loads w B^T T
// (PP:148-154).  The reduction runs over rows (quadrature points x
// components), which are split across workgroups; each workgroup writes an
// FP64 partial tile and a second kernel sums the partials in a fixed order,
// so results are deterministic.
//
// Modes:
//   CR_F32    fp32 operands, v_mfma_f32_16x16x4f32
//   CR_BF16   bf16 operands (converted on load), one bf16 product
//   CR_BF16X3 fp32 operands split on load into hi = bf16(x), lo = bf16(x-hi);
//             x.y ~ hi.hi + hi.lo + lo.hi  (three bf16 MFMAs, fp32 accumulate)
//   CR_F64    fp64 operands, v_mfma_f64_16x16x4f64
// MFMA accumulators are flushed into a second fp32 (fp64 for CR_F64)
// register set every FLUSH rows, which bounds the length of any fp32
// rounding chain; partials are then combined in FP64.
#pragma once
#include "neb_common.hpp"

enum { CR_F32 = 0, CR_BF16 = 1, CR_BF16X3 = 2, CR_F64 = 3 };
constexpr int CR_MAXSEG = 8;
constexpr int CR_TM = 128, CR_TN = 128;

struct CrossSeg {
  const void* x;
  const void* y;
  int64_t ldx, ldy, rows;
  double sign;
};

struct CrossArgs {
  CrossSeg seg[CR_MAXSEG];
  int nseg;
  int64_t m, n;
  const double* roww;  // optional per-row weights (single segment only)
  int symmetric;       // upper tiles only (X == Y)
  int tiles_m, tiles_n, ntiles, splits;
  double* partial;     // [splits][ntiles][TM][TN]
};

template <int MODE> struct CrossPol;
template <> struct CrossPol<CR_F32> {
  using S = float; using frag_t = float; using acc_t = f32x4; using outer_t = float;
  static constexpr int KI = 4, KE = 1, RK = 32, PAD = 4, NPROD = 1, FLUSH = 256;
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct CrossPol<CR_BF16> {
  using S = __bf16; using frag_t = bf16x8; using acc_t = f32x4; using outer_t = float;
  static constexpr int KI = 32, KE = 8, RK = 32, PAD = 8, NPROD = 1, FLUSH = 512;
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
};
template <> struct CrossPol<CR_BF16X3> : CrossPol<CR_BF16> {
  static constexpr int NPROD = 3, FLUSH = 256;
};
template <> struct CrossPol<CR_F64> {
  using S = double; using frag_t = double; using acc_t = f64x4; using outer_t = double;
  static constexpr int KI = 4, KE = 1, RK = 16, PAD = 2, NPROD = 1, FLUSH = 1 << 30;
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
};

template <typename T> __device__ inline double todbl(T v) { return (double)v; }
template <> __device__ inline double todbl<__bf16>(__bf16 v) { return (double)(float)v; }
template <typename T> __device__ inline float toflt(T v) { return (float)v; }
template <> __device__ inline float toflt<__bf16>(__bf16 v) { return (float)v; }

// Transposing tile load: rows [r0, r0+RK) x cols [c0, c0+TM) of a row-major
// source -> LDS hi[col][row] (and lo[col][row] for the split mode).  Thread t
// owns column t % TM and RK/2 consecutive rows, so each global load
// instruction reads TM consecutive elements of one row.
template <int MODE, typename TIN>
__device__ inline void cross_load(typename CrossPol<MODE>::S* hi, typename CrossPol<MODE>::S* lo,
                                  const TIN* __restrict__ src, int64_t ld, int64_t r0, int64_t rows,
                                  int64_t c0, int64_t cols, const double* roww) {
  using Pol = CrossPol<MODE>;
  using S = typename Pol::S;
  constexpr int RK = Pol::RK, LDK = RK + Pol::PAD, HR = RK / 2;
  const int col = threadIdx.x % CR_TM;
  const int rh = (threadIdx.x / CR_TM) * HR;
  const int64_t gc = c0 + col;
  const bool cok = gc < cols;
  S* dh = hi + col * LDK + rh;
  if constexpr (MODE == CR_F64) {
#pragma unroll
    for (int e = 0; e < HR; ++e) {
      const int64_t gr = r0 + rh + e;
      double v = (cok && gr < rows) ? todbl(src[gr * ld + gc]) : 0.0;
      if (roww && gr < rows) v *= roww[gr];
      dh[e] = v;
    }
  } else {
    float v[HR];
#pragma unroll
    for (int e = 0; e < HR; ++e) {
      const int64_t gr = r0 + rh + e;
      v[e] = (cok && gr < rows) ? toflt(src[gr * ld + gc]) : 0.0f;
      if (roww && gr < rows) v[e] = (float)((double)v[e] * roww[gr]);
    }
    if constexpr (MODE == CR_F32) {
#pragma unroll
      for (int e = 0; e < HR; e += 4) *reinterpret_cast<f32x4*>(dh + e) = f32x4{v[e], v[e + 1], v[e + 2], v[e + 3]};
    } else {
#pragma unroll
      for (int e = 0; e < HR; e += 8) {
        bf16x8 h, l;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          __bf16 hq = (__bf16)v[e + q];
          h[q] = hq;
          if constexpr (MODE == CR_BF16X3) l[q] = (__bf16)(v[e + q] - (float)hq);
        }
        *reinterpret_cast<bf16x8*>(dh + e) = h;
        if constexpr (MODE == CR_BF16X3) *reinterpret_cast<bf16x8*>(lo + col * LDK + rh + e) = l;
      }
    }
  }
}

template <int MODE, typename TX, typename TY>
__global__ __launch_bounds__(256) void cross_kernel(CrossArgs a) {
  using Pol = CrossPol<MODE>;
  using S = typename Pol::S;
  using frag_t = typename Pol::frag_t;
  using acc_t = typename Pol::acc_t;
  using outer_t = typename Pol::outer_t;
  constexpr int RK = Pol::RK, LDK = RK + Pol::PAD, KI = Pol::KI, KE = Pol::KE;
  constexpr int NL = (MODE == CR_BF16X3) ? 2 : 1;  // hi (+ lo) tiles
  constexpr int TILE = CR_TM * LDK;

  extern __shared__ __align__(16) unsigned char smem[];
  S* xs = reinterpret_cast<S*>(smem);  // [NL][TM][LDK]
  S* ys = xs + NL * TILE;              // [NL][TN][LDK]

  // tile coordinates (upper-triangular enumeration when symmetric)
  int ti, tj;
  {
    int t = blockIdx.x;
    if (a.symmetric) {
      ti = 0;
      int rowlen = a.tiles_n;
      while (t >= rowlen) { t -= rowlen; ++ti; --rowlen; }
      tj = ti + t;
    } else {
      ti = t / a.tiles_n;
      tj = t % a.tiles_n;
    }
  }
  const int split = blockIdx.y;
  const int64_t i0 = (int64_t)ti * CR_TM, j0 = (int64_t)tj * CR_TN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int fr = lane & 15, kg = (lane >> 4) * KE;

  acc_t inner[4][4];
  outer_t outer[4][4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      inner[i][j] = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int r = 0; r < 4; ++r) outer[i][j][r] = 0;
    }

  for (int s = 0; s < a.nseg; ++s) {
    const CrossSeg sg = a.seg[s];
    int64_t chunk = (sg.rows + a.splits - 1) / a.splits;
    chunk = (chunk + RK - 1) / RK * RK;
    const int64_t rbeg = (int64_t)split * chunk;
    const int64_t rend = rbeg + chunk < sg.rows ? rbeg + chunk : sg.rows;
    if (rbeg >= rend) continue;
    const TX* X = reinterpret_cast<const TX*>(sg.x);
    const TY* Y = reinterpret_cast<const TY*>(sg.y);
    int since = 0;
    for (int64_t r0 = rbeg; r0 < rend; r0 += RK) {
      cross_load<MODE, TX>(xs, xs + TILE, X, sg.ldx, r0, rend, i0, a.m, a.roww);
      cross_load<MODE, TY>(ys, ys + TILE, Y, sg.ldy, r0, rend, j0, a.n, nullptr);
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < RK; kk += KI) {
        frag_t xh[4], yh[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) xh[i] = *reinterpret_cast<const frag_t*>(xs + (wi + i * 16 + fr) * LDK + kk + kg);
#pragma unroll
        for (int j = 0; j < 4; ++j) yh[j] = *reinterpret_cast<const frag_t*>(ys + (wj + j * 16 + fr) * LDK + kk + kg);
        if constexpr (MODE == CR_BF16X3) {
          frag_t xl[4], yl[4];
#pragma unroll
          for (int i = 0; i < 4; ++i) xl[i] = *reinterpret_cast<const frag_t*>(xs + TILE + (wi + i * 16 + fr) * LDK + kk + kg);
#pragma unroll
          for (int j = 0; j < 4; ++j) yl[j] = *reinterpret_cast<const frag_t*>(ys + TILE + (wj + j * 16 + fr) * LDK + kk + kg);
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              inner[i][j] = Pol::mfma(xl[i], yh[j], inner[i][j]);
              inner[i][j] = Pol::mfma(xh[i], yl[j], inner[i][j]);
              inner[i][j] = Pol::mfma(xh[i], yh[j], inner[i][j]);
            }
        } else {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) inner[i][j] = Pol::mfma(xh[i], yh[j], inner[i][j]);
        }
      }
      __syncthreads();
      since += RK;
      if (since >= Pol::FLUSH || r0 + RK >= rend) {
        const outer_t sgn = (outer_t)sg.sign;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) outer[i][j][r] += sgn * (outer_t)inner[i][j][r];
            inner[i][j] = acc_t{0, 0, 0, 0};
          }
        since = 0;
      }
    }
  }

  double* out = a.partial + ((int64_t)split * a.ntiles + blockIdx.x) * (CR_TM * CR_TN);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int li = wi + i * 16 + Pol::drow(lane, r);
        const int lj = wj + j * 16 + fr;
        out[li * CR_TN + lj] = (double)outer[i][j][r];
      }
}

// G[i,j] = beta*G[i,j] + scale * sum_split partial  (mirrored when symmetric)
__global__ __launch_bounds__(256) void cross_reduce_kernel(const double* __restrict__ partial, int splits,
                                                            int ntiles, int tiles_n, int symmetric, int64_t m,
                                                            int64_t n, double scale, double beta,
                                                            double* __restrict__ G, int64_t ldg) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= m * n) return;
  int64_t i = idx / n, j = idx % n;
  if (symmetric && i > j) {  // read the upper triangle: exact symmetry
    const int64_t t = i; i = j; j = t;
  }
  const int64_t ti = i / CR_TM, tj = j / CR_TN, li = i % CR_TM, lj = j % CR_TN;
  int64_t tile;
  if (symmetric) tile = ti * tiles_n - ti * (ti - 1) / 2 + (tj - ti);
  else tile = ti * tiles_n + tj;
  double s = 0.0;
  const double* p = partial + tile * (CR_TM * CR_TN) + li * CR_TN + lj;
  const int64_t stride = (int64_t)ntiles * CR_TM * CR_TN;
  for (int q = 0; q < splits; ++q) s += p[q * stride];
  double* g = G + (idx / n) * ldg + (idx % n);
  *g = (beta == 0.0 ? 0.0 : beta * *g) + scale * s;
}

template <int MODE>
constexpr size_t cross_lds_bytes() {
  using Pol = CrossPol<MODE>;
  constexpr int NL = (MODE == CR_BF16X3) ? 2 : 1;
  return sizeof(typename Pol::S) * (size_t)(Pol::RK + Pol::PAD) * CR_TM * NL * 2;
}
