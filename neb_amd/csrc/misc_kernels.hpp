// Elementwise / reduction kernels around the contractions: periodic feature
// jets, column sums (bias gradients), sum of squares (gradient norm), Adam,
// and weight staging (precision conversion, transpose, zero padding).
#pragma once
#include <type_traits>

#include "neb_common.hpp"

// Periodic sin/cos feature jets (HB2P:566-589, periodic branch; HB2P:428-441).
// Feature f < 2*F*D: coordinate a = f / (2F), harmonic h = (f % 2F)/2 + 1,
// even f -> sin, odd f -> cos of phase = pi*h*(xn+1), xn = 2(x-lo)/(hi-lo)-1.
// Features in [2FD, ldf) are zero padding.  Output packed [C][n][ldf].
template <int D, typename TX, typename S>
__global__ __launch_bounds__(256) void feature_kernel(const TX* __restrict__ x, int64_t n, int freqs,
                                                      double lo0, double lo1, double lo2, double hi0,
                                                      double hi1, double hi2, S* __restrict__ out, int64_t ldf) {
  using L = Jet2Layout<D>;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * ldf) return;
  const int64_t p = idx / ldf;
  const int f = (int)(idx % ldf);
  const int64_t cs = n * ldf;
  double ch[L::C];
#pragma unroll
  for (int c = 0; c < L::C; ++c) ch[c] = 0.0;
  if (f < 2 * freqs * D) {
    const int a = f / (2 * freqs);
    const int h = (f % (2 * freqs)) / 2 + 1;
    const double lo = a == 0 ? lo0 : (a == 1 ? lo1 : lo2);
    const double hi = a == 0 ? hi0 : (a == 1 ? hi1 : hi2);
    const double sc = 2.0 / (hi - lo);
    const double xn = sc * ((double)x[p * D + a] - lo) - 1.0;
    const double k = M_PI * (double)h;
    const double ph = k * (xn + 1.0);
    const double dp = k * sc;
    const double s = sin(ph), c = cos(ph);
    const bool is_cos = f & 1;
    const double v = is_cos ? c : s;
    const double g = is_cos ? -s * dp : c * dp;
    ch[0] = v;
    ch[1 + a] = g;
    // diagonal pair (a,a): index in PAIRS order
    const int diag = D == 2 ? (a == 0 ? 0 : 2) : (a == 0 ? 0 : (a == 1 ? 3 : 5));
    ch[1 + D + diag] = -v * dp * dp;
  }
#pragma unroll
  for (int c = 0; c < L::C; ++c) out[c * cs + p * ldf + f] = cvt<S>(ch[c]);
}

// Column sums in FP64: partial[split][col] over the split's row range
// (blockIdx.x) for one slab of colsum_slab_groups(cols) 16-byte column
// groups (blockIdx.y).  Each thread sums 16-byte column vectors down a
// strided set of rows; row lanes are then combined through LDS.
template <typename S> __host__ __device__ inline int64_t colsum_slab_groups(int64_t cols) {
  const int64_t ngroups = (cols + 16 / (int64_t)sizeof(S) - 1) / (16 / (int64_t)sizeof(S));
  return ngroups < 32 ? ngroups : 32;
}

template <typename S>
__global__ __launch_bounds__(256) void colsum_kernel(const S* __restrict__ x, int64_t ld, int64_t rows,
                                                     int64_t cols, int64_t chunk, double* __restrict__ partial) {
  constexpr int VE = 16 / sizeof(S);
  typedef S vs __attribute__((ext_vector_type(VE)));
  __shared__ double red[256 * VE];
  const int64_t rbeg = (int64_t)blockIdx.x * chunk;
  const int64_t rend = rbeg + chunk < rows ? rbeg + chunk : rows;
  const int64_t ngroups = (cols + VE - 1) / VE;
  const int CG = (int)colsum_slab_groups<S>(cols);
  const int RL = 256 / CG;
  const int cgl = threadIdx.x % CG, rl = threadIdx.x / CG;
  {
    const int64_t g0 = (int64_t)blockIdx.y * CG;
    double acc[VE];
#pragma unroll
    for (int e = 0; e < VE; ++e) acc[e] = 0.0;
    const int64_t col = (g0 + cgl) * VE;
    if (rl < RL && g0 + cgl < ngroups) {
      if (col + VE <= cols) {  // 8 rows of 16-byte loads in flight, added in row order
        int64_t r = rbeg + rl;
        for (; r + 7 * RL < rend; r += 8 * RL) {
          vs v[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const vs*>(x + (r + u * RL) * ld + col);
#pragma unroll
          for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int e = 0; e < VE; ++e) acc[e] += todbl_any(v[u][e]);
        }
        for (; r < rend; r += RL) {
          const vs v = *reinterpret_cast<const vs*>(x + r * ld + col);
#pragma unroll
          for (int e = 0; e < VE; ++e) acc[e] += todbl_any(v[e]);
        }
      } else {
        for (int64_t r = rbeg + rl; r < rend; r += RL) {
          const S* p = x + r * ld + col;
#pragma unroll
          for (int e = 0; e < VE; ++e)
            if (col + e < cols) acc[e] += todbl_any(p[e]);
        }
      }
    }
    if (rl < RL)
#pragma unroll
      for (int e = 0; e < VE; ++e) red[(rl * CG + cgl) * VE + e] = acc[e];
    __syncthreads();
    for (int t = threadIdx.x; t < CG * VE; t += 256) {
      double sacc = 0.0;
      for (int q = 0; q < RL; ++q) sacc += red[q * CG * VE + t];
      const int64_t c = g0 * VE + t;
      if (c < cols) partial[(int64_t)blockIdx.x * cols + c] = sacc;
    }
  }
}

// Mode-diagonal crosses: out[m][j] = sum_s sign_s sum_r x_s[r][m] y_s[r][m h + j]
// (the diagonal 1 x h blocks of sum_s sign_s X_s^T Y_s for a [B, B h] result:
// each mode's output-weight products against its own hidden features).
// Thread = 8 consecutive columns (one mode's) over the rows of one chunk
// (blockIdx.x), FP64 accumulation; partial[chunk][B h] is then summed over
// the chunks in fixed order (sum_splits_kernel).
constexpr int MC_MAXSEG = 16;
struct ModeCrossArgs {
  const void* x[MC_MAXSEG];
  const void* y[MC_MAXSEG];
  int64_t ldx[MC_MAXSEG], ldy[MC_MAXSEG], rows[MC_MAXSEG];
  double sign[MC_MAXSEG];
  int nseg;
  int64_t B, h, chunk;
  double scale;
  double* partial;
};

// 8 consecutive elements (16-byte aligned) as FP64
template <typename T> __device__ inline void ld8_f64(const T* src, double* v) {
  constexpr int PV = 16 / sizeof(T);
  typedef T pv __attribute__((ext_vector_type(PV)));
#pragma unroll
  for (int c = 0; c < 8; c += PV) {
    const pv t = *reinterpret_cast<const pv*>(src + c);
#pragma unroll
    for (int e = 0; e < PV; ++e) v[c + e] = todbl_any(t[e]);
  }
}

template <typename TX, typename TY>
__global__ __launch_bounds__(256) void mode_cross_kernel(ModeCrossArgs a) {
  const int64_t ncol = a.B * a.h;
  const int64_t g = (int64_t)blockIdx.y * blockDim.x + threadIdx.x;
  if (g * 8 >= ncol) return;
  const int64_t c0 = g * 8, m = c0 / a.h;
  const int64_t r0 = (int64_t)blockIdx.x * a.chunk;
  double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int s = 0; s < a.nseg; ++s) {
    const TX* xs = reinterpret_cast<const TX*>(a.x[s]);
    const TY* ys = reinterpret_cast<const TY*>(a.y[s]);
    const int64_t r1 = r0 + a.chunk < a.rows[s] ? r0 + a.chunk : a.rows[s];
    double sa[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t r = r0; r < r1; ++r) {
      const double xv = todbl_any(xs[r * a.ldx[s] + m]);
      double yv[8];
      ld8_f64<TY>(ys + r * a.ldy[s] + c0, yv);
#pragma unroll
      for (int e = 0; e < 8; ++e) sa[e] = fma(xv, yv[e], sa[e]);
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) acc[e] = fma(a.sign[s], sa[e], acc[e]);
  }
  double* o = a.partial + (int64_t)blockIdx.x * ncol + c0;
#pragma unroll
  for (int e = 0; e < 8; ++e) o[e] = a.scale * acc[e];
}

// out[c] = beta out[c] + sum_s partial[s][c] in a fixed order: SS_COLS
// columns per workgroup, SS_LANES split lanes per column, lane l summing
// splits l, l + SS_LANES, ... in order (8 loads in flight), then the lanes
// combined in lane order through LDS.  The order depends only on `splits`.
constexpr int SS_COLS = 16, SS_LANES = 16;
inline dim3 sum_splits_grid(int64_t cols) { return dim3((unsigned)((cols + SS_COLS - 1) / SS_COLS)); }

__global__ __launch_bounds__(256) void sum_splits_kernel(const double* __restrict__ partial, int splits,
                                                         int64_t cols, double* __restrict__ out, double beta) {
  static_assert(SS_COLS * SS_LANES == 256, "one thread per (column, lane)");
  __shared__ double red[SS_LANES][SS_COLS];
  const int cl = threadIdx.x % SS_COLS, lane = threadIdx.x / SS_COLS;
  const int64_t c = (int64_t)blockIdx.x * SS_COLS + cl;
  double s = 0.0;
  if (c < cols) {
    int q = lane;
    for (; q + 7 * SS_LANES < splits; q += 8 * SS_LANES) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = partial[(int64_t)(q + u * SS_LANES) * cols + c];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += t[u];
    }
    for (; q < splits; q += SS_LANES) s += partial[(int64_t)q * cols + c];
  }
  red[lane][cl] = s;
  __syncthreads();
  if (lane == 0 && c < cols) {
    double t = red[0][cl];
#pragma unroll
    for (int l = 1; l < SS_LANES; ++l) t += red[l][cl];
    out[c] = (beta == 0.0 ? 0.0 : beta * out[c]) + t;
  }
}

// Sum of squares in FP64 (gradient norm for the Adam clip, NCC:157).
template <typename T>
__global__ __launch_bounds__(256) void sumsq_kernel(const T* __restrict__ x, int64_t n, double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = todbl_any(x[i]);
    s += v * v;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam with bias correction and a precomputed clip scale (NCC:156-171).
template <typename T>
__global__ __launch_bounds__(256) void adam_kernel(T* __restrict__ p, const T* __restrict__ g, T* __restrict__ m,
                                                   T* __restrict__ v, int64_t n, double lr, double scale, double b1,
                                                   double b2, double eps, double c1, double c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double gi = (double)g[i] * scale;
  const double mi = b1 * (double)m[i] + (1.0 - b1) * gi;
  const double vi = b2 * (double)v[i] + (1.0 - b2) * gi * gi;
  m[i] = (T)mi;
  v[i] = (T)vi;
  p[i] = (T)((double)p[i] - lr * (mi / c1) / (sqrt(vi / c2) + eps));
}

// Adam with the gates of the step evaluated on the device: the clip scale
// min(1, clip / |g|) comes from the squared gradient norm *norm2, and nothing
// is written when |g| is zero or nonfinite or *gate (an earlier gate of the
// step, nullable) is nonzero (NCC:156-171 clip and update).
template <typename T>
__global__ __launch_bounds__(256) void adam_gated_kernel(T* __restrict__ p, const T* __restrict__ g, T* __restrict__ m,
                                                         T* __restrict__ v, int64_t n, double lr, double clip,
                                                         double b1, double b2, double eps, double c1, double c2,
                                                         const double* __restrict__ norm2,
                                                         const int* __restrict__ gate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double nrm = sqrt(*norm2);
  if ((gate && *gate != 0) || !(nrm > 0.0) || !isfinite(nrm)) return;
  const double scale = fmin(1.0, clip / fmax(nrm, 1e-30));
  const double gi = (double)g[i] * scale;
  const double mi = b1 * (double)m[i] + (1.0 - b1) * gi;
  const double vi = b2 * (double)v[i] + (1.0 - b2) * gi * gi;
  m[i] = (T)mi;
  v[i] = (T)vi;
  p[i] = (T)((double)p[i] - lr * (mi / c1) / (sqrt(vi / c2) + eps));
}

// dst[r][c] (dst_rows x dst_cols, zero outside the source) = src[r][c] or
// src[c][r] when transposing; converts the master precision to the MFMA
// operand precision.
// Split of a matrix into bf16 planes: hi = bf16(x), lo = bf16(x - hi)
// (the difference taken in the source precision, exact for fp32 sources).
template <typename TS>
__global__ __launch_bounds__(256) void split_kernel(const TS* __restrict__ src, int64_t rows, int64_t cols,
                                                    int64_t lds, __bf16* __restrict__ hi, __bf16* __restrict__ lo,
                                                    int64_t ldd) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * cols) return;
  const int64_t r = idx / cols, c = idx % cols;
  const double x = todbl_any(src[r * lds + c]);
  const __bf16 h = (__bf16)(float)x;
  hi[r * ldd + c] = h;
  lo[r * ldd + c] = (__bf16)(float)(x - (double)(float)h);
}

// Weights of one block of B per-mode networks as a single layer (the mode
// block as a grouped GEMM): from the mode-stacked master S [K][out][in]
// (row-major per mode), modes m0 .. m0+B-1:
//   diag = 0 (the first layer, every mode reads the same input features):
//     rows m*out + o, columns i           -> S[m0+m][o][i]
//   diag = 1 (hidden and output layers, every mode reads its own inputs):
//     rows m*out + o, columns m'*in + i   -> S[m0+m][o][i] if m == m' else 0
// written as [B*out][ld] (columns beyond the width zero) in the destination
// precision, optionally with the split lo plane (bf16 only) and the transpose
// [width][ldt] (width = in or B*in).
template <typename TS, typename TD>
__global__ __launch_bounds__(256) void stage_modes_kernel(const TS* __restrict__ S, int64_t out, int64_t in,
                                                          int64_t m0, int64_t B, int diag, TD* __restrict__ hi,
                                                          TD* __restrict__ lo, int64_t ld, TD* __restrict__ wt,
                                                          int64_t ldt) {
  const int64_t rows = B * out, cols = diag ? B * in : in;
  const int64_t nw = rows * ld;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (diag == 2) {  // the diagonal blocks only, mode by mode (the rest of hi / lo / wt is left as it is)
    const int64_t nd = B * out * in;
    if (idx < nd) {
      const int64_t m = idx / (out * in), o = (idx / in) % out, i = idx % in;
      const double x = todbl_any(S[((m0 + m) * out + o) * in + i]);
      const TD h = cvt<TD>(x);
      const int64_t at = (m * out + o) * ld + m * in + i;
      hi[at] = h;
      if (lo) lo[at] = cvt<TD>(x - todbl_any(h));
      return;
    }
    const int64_t it = idx - nd;  // wt[m in + i][m out + o], o fastest
    if (!wt || it >= nd) return;
    const int64_t m = it / (in * out), i = (it / out) % in, o = it % out;
    wt[(m * in + i) * ldt + m * out + o] = cvt<TD>(todbl_any(S[((m0 + m) * out + o) * in + i]));
    return;
  }
  auto value = [&](int64_t r, int64_t c) -> double {
    if (r >= rows || c >= cols) return 0.0;
    const int64_t m = r / out, o = r % out;
    int64_t i = c;
    if (diag) {
      if (c / in != m) return 0.0;
      i = c % in;
    }
    return todbl_any(S[((m0 + m) * out + o) * in + i]);
  };
  if (idx < nw) {
    const double x = value(idx / ld, idx % ld);
    const TD h = cvt<TD>(x);
    hi[idx] = h;
    if (lo) lo[idx] = cvt<TD>(x - todbl_any(h));
    return;
  }
  const int64_t it = idx - nw;
  if (!wt || it >= cols * ldt) return;
  wt[it] = cvt<TD>(value(it % ldt, it / ldt));
}

// Transpose of the envelope product (jet_product_reverse, VJ2:51-55) for
// every (point, output) pair: adj [C][n][lda] (math precision) -> out
// [C][n][ldo] in the operand precision of the output-layer reverse.
template <typename SO, typename M, int D>
__global__ __launch_bounds__(256) void envelope_reverse_kernel(const M* __restrict__ env, const M* __restrict__ adj,
                                                               int64_t lda, int64_t n, int64_t nout,
                                                               SO* __restrict__ out, int64_t ldo) {
  constexpr int C = Jet2Layout<D>::C;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * nout) return;
  const int64_t p = idx / nout, o = idx % nout;
  M e[C], a[C], r[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    e[c] = env[c * n + p];
    a[c] = adj[(c * n + p) * lda + o];
  }
  jet_product_reverse<D, M>(e, a, r);
#pragma unroll
  for (int c = 0; c < C; ++c) out[(c * n + p) * ldo + o] = cvt<SO>(r[c]);
}

// Every layer's bf16 operand forms in one launch (see neb_stage_net): a flat
// index over, first, the [rows][ld] planes of all layers (hi, and lo when
// requested), then the [cols][ldt] transposes; padding columns are zeros.
constexpr int STAGE_NET_MAX = 16;
struct StageNetArgs {
  int nl;
  int64_t off[STAGE_NET_MAX], rows[STAGE_NET_MAX], cols[STAGE_NET_MAX];
  __bf16* hi[STAGE_NET_MAX];
  __bf16* lo[STAGE_NET_MAX];
  __bf16* wt[STAGE_NET_MAX];
  int64_t ld[STAGE_NET_MAX], ldt[STAGE_NET_MAX];
  int64_t start_w[STAGE_NET_MAX + 1], start_t[STAGE_NET_MAX + 1];
};

template <typename TS>
__global__ __launch_bounds__(256) void stage_net_kernel(const TS* __restrict__ master, StageNetArgs a) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nw = a.start_w[a.nl];
  if (idx < nw) {
    int l = 0;
    while (idx >= a.start_w[l + 1]) ++l;
    const int64_t e = idx - a.start_w[l], r = e / a.ld[l], c = e % a.ld[l];
    // the rounding of split_kernel: hi = bf16(x), lo = bf16(x - hi) with the
    // difference taken in double
    const double x = c < a.cols[l] ? (double)master[a.off[l] + r * a.cols[l] + c] : 0.0;
    const __bf16 h = (__bf16)(float)x;
    a.hi[l][e] = h;
    if (a.lo[l]) a.lo[l][e] = (__bf16)(float)(x - (double)(float)h);
    return;
  }
  const int64_t it = idx - nw;
  if (it >= a.start_t[a.nl]) return;
  int l = 0;
  while (it >= a.start_t[l + 1]) ++l;
  const int64_t e = it - a.start_t[l], i = e / a.ldt[l], o = e % a.ldt[l];
  a.wt[l][e] = o < a.rows[l] ? (__bf16)(float)master[a.off[l] + o * a.cols[l] + i] : (__bf16)0.0f;
}

// Join of bf16 hi / lo planes into one fp32 or fp64 matrix: dst = hi + lo
// (exact in fp32: lo holds the rounding residual of hi).  Each thread joins
// eight consecutive columns of one row.
template <typename TD>
__global__ __launch_bounds__(256) void join_kernel(const __bf16* __restrict__ hi, const __bf16* __restrict__ lo,
                                                   int64_t rows, int64_t cols, int64_t ld, TD* __restrict__ dst,
                                                   int64_t ldd) {
  const int64_t per_row = (cols + 7) / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * per_row) return;
  const int64_t r = idx / per_row, c0 = (idx % per_row) * 8;
  const __bf16* h = hi + r * ld;
  const __bf16* l = lo + r * ld;
  TD* d = dst + r * ldd;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t c = c0 + q;
    if (c < cols) d[c] = (TD)((float)h[c] + (float)l[c]);
  }
}

// Column centering x[r][c] -= mean[c] (see neb_center_columns).  Split
// planes (lo not null): hi + lo - mean in FP64, split again like
// split_kernel.  Each thread handles eight consecutive columns of one row.
template <typename T>
__global__ __launch_bounds__(256) void center_kernel(T* __restrict__ x, __bf16* __restrict__ lo, int64_t rows,
                                                     int64_t cols, int64_t ld, const double* __restrict__ mean) {
  const int64_t per_row = (cols + 7) / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * per_row) return;
  const int64_t r = idx / per_row, c0 = (idx % per_row) * 8;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int64_t c = c0 + q;
    if (c >= cols) break;
    T* p = x + r * ld + c;
    if constexpr (std::is_same<T, __bf16>::value) {
      if (lo) {
        __bf16* pl = lo + r * ld + c;
        const double v = (double)(float)*p + (double)(float)*pl - mean[c];
        const __bf16 h = (__bf16)(float)v;
        *p = h;
        *pl = (__bf16)(float)(v - (double)(float)h);
      } else {
        *p = (__bf16)(float)((double)(float)*p - mean[c]);
      }
    } else {
      *p = (T)((double)*p - mean[c]);
    }
  }
}

template <typename TS, typename TD>
__global__ __launch_bounds__(256) void stage_kernel(const TS* __restrict__ src, int64_t rows, int64_t cols,
                                                    int64_t lds, TD* __restrict__ dst, int64_t ldd,
                                                    int64_t drows, int64_t dcols, int transpose, double scale) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= drows * dcols) return;
  const int64_t r = idx / dcols, c = idx % dcols;
  double v = 0.0;
  if (transpose) {
    if (c < rows && r < cols) v = todbl_any(src[c * lds + r]);
  } else {
    if (r < rows && c < cols) v = todbl_any(src[r * lds + c]);
  }
  dst[r * ldd + c] = cvt<TD>(v * scale);
}

// Channels of a packed jet tensor side by side, point-major (see
// neb_stack_channels): dst[p][i * width + j] = src[ch[i]][p][j] for the
// hi plane (and the lo plane when given), eight bf16 per thread.
__global__ __launch_bounds__(256) void stack_channels_kernel(const __bf16* __restrict__ src_hi,
                                                             const __bf16* __restrict__ src_lo, int64_t n,
                                                             int64_t ld, int64_t width, ChanMap map,
                                                             __bf16* __restrict__ dst_hi, __bf16* __restrict__ dst_lo,
                                                             int64_t ldd) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const int64_t per = width / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * map.n * per) return;
  const int64_t j = (idx % per) * 8, q = idx / per;
  const int i = (int)(q % map.n);
  const int64_t p = q / map.n;
  int c = map.ch[0];
#pragma unroll
  for (int t = 1; t < STACK_MAXCH; ++t)
    if (i == t) c = map.ch[t];
  const int64_t s = ((int64_t)c * n + p) * ld + j, d = p * ldd + (int64_t)i * width + j;
  *reinterpret_cast<u32x4*>(dst_hi + d) = *reinterpret_cast<const u32x4*>(src_hi + s);
  if (src_lo) *reinterpret_cast<u32x4*>(dst_lo + d) = *reinterpret_cast<const u32x4*>(src_lo + s);
}
