// Elementwise / reduction kernels around the contractions: periodic feature
// jets, column sums (bias gradients), sum of squares (gradient norm), Adam,
// and weight staging (precision conversion, transpose, zero padding).
#pragma once
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

// Column sums in FP64: partial[s][col] over a row range per split.
template <typename S>
__global__ __launch_bounds__(256) void colsum_kernel(const S* __restrict__ x, int64_t ld, int64_t rows,
                                                     int64_t cols, int64_t chunk, double* __restrict__ partial) {
  __shared__ double red[4][64];
  const int64_t col = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int rl = threadIdx.x >> 6;
  const int64_t rbeg = (int64_t)blockIdx.y * chunk;
  int64_t rend = rbeg + chunk;
  if (rend > rows) rend = rows;
  double s = 0.0;
  if (col < cols)
    for (int64_t r = rbeg + rl; r < rend; r += 4) s += todbl_any(x[r * ld + col]);
  red[rl][threadIdx.x & 63] = s;
  __syncthreads();
  if (rl == 0 && col < cols)
    partial[(int64_t)blockIdx.y * cols + col] = red[0][threadIdx.x] + red[1][threadIdx.x] +
                                                red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[c] = sum_s partial[s][c] in fixed order
__global__ __launch_bounds__(256) void sum_splits_kernel(const double* __restrict__ partial, int splits,
                                                         int64_t cols, double* __restrict__ out, double beta) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  double s = 0.0;
  for (int q = 0; q < splits; ++q) s += partial[(int64_t)q * cols + c];
  out[c] = (beta == 0.0 ? 0.0 : beta * out[c]) + s;
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

// dst[r][c] (dst_rows x dst_cols, zero outside the source) = src[r][c] or
// src[c][r] when transposing; converts the master precision to the MFMA
// operand precision.
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
