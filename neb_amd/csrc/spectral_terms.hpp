// Elementwise and reduction terms of the FP64 spectral objective around the
// SPD inverse and the X S X product (MB/velocity_trace_jet2.py:107-117):
//   M_r = sym(M) + ridge I                                  (sym_shift_kernel)
//   X <- sym(X), S_s = sym(S),
//   L = <X, S_s> + lam ||M - I||_F^2,  cond = ||M_r||_F ||X||_F (spectral_terms_kernel
//                                                            + spectral_final_kernel)
//   dL/dM = -sym(C) + 2 lam (M - I),  C = X S_s X           (spectral_adjoint_kernel)
// sym(A) = (A + A^T) / 2.  Each kernel works on 32 x 32 tile pairs (bi, bj),
// bi <= bj, and its mirror (bj, bi): both tiles are read through LDS before
// either is written, so outputs may alias inputs.  Reductions are
// deterministic: one partial per workgroup, summed in a fixed order.
#pragma once
#include "neb_common.hpp"

constexpr int SYM_T = 32;

// Loads tiles (bi, bj) and (bj, bi) of A (identity-free, zero outside K) into
// t[0] and t[1] ([row][col], padded against bank conflicts); 256 threads as
// 32 columns x 8 rows.
__device__ inline void sym_load_pair(const double* __restrict__ A, int64_t lda, int64_t K, int bi, int bj,
                                     double (*t)[SYM_T][SYM_T + 1]) {
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = ty + 8 * u;
    const int64_t i0 = (int64_t)bi * SYM_T + r, j0 = (int64_t)bj * SYM_T + tx;
    const int64_t i1 = (int64_t)bj * SYM_T + r, j1 = (int64_t)bi * SYM_T + tx;
    t[0][r][tx] = (i0 < K && j0 < K) ? A[i0 * lda + j0] : 0.0;
    t[1][r][tx] = (i1 < K && j1 < K) ? A[i1 * lda + j1] : 0.0;
  }
}

// B = sym(A) + shift I (B may alias A).
__global__ __launch_bounds__(256) void sym_shift_kernel(int64_t K, const double* A, int64_t lda, double shift,
                                                        double* B, int64_t ldb) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  __shared__ double t[2][SYM_T][SYM_T + 1];
  sym_load_pair(A, lda, K, bi, bj, t);
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = ty + 8 * u;
    const int64_t i0 = (int64_t)bi * SYM_T + r, j0 = (int64_t)bj * SYM_T + tx;
    const int64_t i1 = (int64_t)bj * SYM_T + r, j1 = (int64_t)bi * SYM_T + tx;
    if (i0 < K && j0 < K) B[i0 * ldb + j0] = 0.5 * (t[0][r][tx] + t[1][tx][r]) + (i0 == j0 ? shift : 0.0);
    if (bi != bj && i1 < K && j1 < K) B[i1 * ldb + j1] = 0.5 * (t[1][r][tx] + t[0][tx][r]);
  }
}

// X <- sym(X), Ss = sym(S); partial sums per workgroup (slot bj + nt * bi):
//   [0] <X, Ss>, [1] ||M - I||^2, [2] ||sym(M) + ridge I||^2, [3] ||X||^2.
__global__ __launch_bounds__(256) void spectral_terms_kernel(int64_t K, const double* M, int64_t ldm,
                                                             const double* S, int64_t lds, double* X, int64_t ldx,
                                                             double* Ss, int64_t ldss, double ridge,
                                                             double* __restrict__ part) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  const int slot = bj + gridDim.x * bi;
  __shared__ double tx_[2][SYM_T][SYM_T + 1], ts[2][SYM_T][SYM_T + 1], tm[2][SYM_T][SYM_T + 1];
  __shared__ double red[4][256];
  const int t = threadIdx.x, tx = t & 31, ty = t >> 5;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  if (bi <= bj) {
    sym_load_pair(X, ldx, K, bi, bj, tx_);
    sym_load_pair(S, lds, K, bi, bj, ts);
    sym_load_pair(M, ldm, K, bi, bj, tm);
    __syncthreads();
    const bool diag = bi == bj;
    const double mult = diag ? 1.0 : 2.0;  // an off-diagonal pair holds each symmetric value twice
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int r = ty + 8 * u;
      const int64_t i0 = (int64_t)bi * SYM_T + r, j0 = (int64_t)bj * SYM_T + tx;
      const int64_t i1 = (int64_t)bj * SYM_T + r, j1 = (int64_t)bi * SYM_T + tx;
      if (i0 < K && j0 < K) {
        const double xs = 0.5 * (tx_[0][r][tx] + tx_[1][tx][r]);
        const double ss = 0.5 * (ts[0][r][tx] + ts[1][tx][r]);
        const double mr = 0.5 * (tm[0][r][tx] + tm[1][tx][r]) + (i0 == j0 ? ridge : 0.0);
        const double d0 = tm[0][r][tx] - (i0 == j0 ? 1.0 : 0.0);
        acc[0] += mult * xs * ss;
        acc[1] += d0 * d0;
        acc[2] += mult * mr * mr;
        acc[3] += mult * xs * xs;
        X[i0 * ldx + j0] = xs;
        Ss[i0 * ldss + j0] = ss;
      }
      if (!diag && i1 < K && j1 < K) {
        const double d1 = tm[1][r][tx];
        acc[1] += d1 * d1;
        X[i1 * ldx + j1] = 0.5 * (tx_[1][r][tx] + tx_[0][tx][r]);
        Ss[i1 * ldss + j1] = 0.5 * (ts[1][r][tx] + ts[0][tx][r]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + w];
    }
    __syncthreads();
  }
  if (t < 4) part[(int64_t)slot * 4 + t] = red[t][0];
}

// Sums the nparts partials in a fixed order; loss = [0] + lam [1],
// cond = sqrt([2]) sqrt([3]); gate = *info when nonzero (first nonpositive
// pivot of the inverse, nullable), else -1 for a nonfinite or too large
// condition estimate, else 0.
__global__ __launch_bounds__(256) void spectral_final_kernel(const double* __restrict__ part, int64_t nparts,
                                                             const int* __restrict__ info, double lam,
                                                             double cond_limit, double* __restrict__ loss,
                                                             int* __restrict__ gate) {
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t p = t; p < nparts; p += 256) {
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] += part[p * 4 + q];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) red[q][t] = acc[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[q][t] += red[q][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    *loss = red[0][0] + lam * red[1][0];
    const double cond = sqrt(red[2][0]) * sqrt(red[3][0]);
    const int bad = info ? *info : 0;
    *gate = bad != 0 ? bad : ((isfinite(cond) && cond <= cond_limit) ? 0 : -1);
  }
}

// dM = -sym(C) + 2 lam (M - I)  (dM may alias C).
__global__ __launch_bounds__(256) void spectral_adjoint_kernel(int64_t K, const double* C, int64_t ldc,
                                                               const double* M, int64_t ldm, double lam, double* dM,
                                                               int64_t ldd) {
  const int bi = blockIdx.y, bj = blockIdx.x;
  if (bi > bj) return;
  __shared__ double t[2][SYM_T][SYM_T + 1];
  sym_load_pair(C, ldc, K, bi, bj, t);
  __syncthreads();
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int r = ty + 8 * u;
    const int64_t i0 = (int64_t)bi * SYM_T + r, j0 = (int64_t)bj * SYM_T + tx;
    const int64_t i1 = (int64_t)bj * SYM_T + r, j1 = (int64_t)bi * SYM_T + tx;
    if (i0 < K && j0 < K)
      dM[i0 * ldd + j0] = -0.5 * (t[0][r][tx] + t[1][tx][r]) + 2.0 * lam * (M[i0 * ldm + j0] - (i0 == j0 ? 1.0 : 0.0));
    if (bi != bj && i1 < K && j1 < K)
      dM[i1 * ldd + j1] = -0.5 * (t[1][r][tx] + t[0][tx][r]) + 2.0 * lam * M[i1 * ldm + j1];
  }
}
