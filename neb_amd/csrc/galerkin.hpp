// Reduced-order (Galerkin) right-hand side of the 2D periodic rollout on the
// network basis (experiments/periodic/paper_rollout.py:37-41):
//   u = gu a, v = gv a, omega = gw a          (per point, K-long dot products)
//   rhs = gu^T (w omega v) - gv^T (w omega u)  (column sums over the points)
// Two passes over the [n][K] field matrices, both HBM-bound: the point pass
// reads each point's three rows once (one wave per point, lanes along K,
// the coefficients in LDS), the column pass reads gu and gv once more with
// one thread per column and a row range per block, partial sums in FP64
// reduced in a fixed order (sum_splits_kernel).
#pragma once
#include <hip/hip_runtime.h>

#include "neb_common.hpp"

template <typename T>
__global__ __launch_bounds__(256) void rhs2d_points_kernel(const T* __restrict__ gu, const T* __restrict__ gv,
                                                           const T* __restrict__ gw, int64_t ld, int64_t n, int64_t K,
                                                           const double* __restrict__ a, double w,
                                                           double* __restrict__ r1, double* __restrict__ r2) {
  extern __shared__ double sa[];
  for (int64_t i = threadIdx.x; i < K; i += blockDim.x) sa[i] = a[i];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t p = (int64_t)blockIdx.x * (blockDim.x >> 6) + wave;
  if (p >= n) return;
  const T* ru = gu + p * ld;
  const T* rv = gv + p * ld;
  const T* rw = gw + p * ld;
  double su = 0.0, sv = 0.0, so = 0.0;
  for (int64_t k = lane; k < K; k += 64) {
    const double ak = sa[k];
    su += todbl_any(ru[k]) * ak;
    sv += todbl_any(rv[k]) * ak;
    so += todbl_any(rw[k]) * ak;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    su += __shfl_xor(su, off);
    sv += __shfl_xor(sv, off);
    so += __shfl_xor(so, off);
  }
  if (lane == 0) {
    r1[p] = w * so * sv;
    r2[p] = w * so * su;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void rhs2d_cols_kernel(const T* __restrict__ gu, const T* __restrict__ gv,
                                                         int64_t ld, int64_t n, int64_t K, int64_t chunk,
                                                         const double* __restrict__ r1,
                                                         const double* __restrict__ r2,
                                                         double* __restrict__ partial) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.y * chunk;
  const int64_t p1 = p0 + chunk < n ? p0 + chunk : n;
  if (k >= K) return;
  double s = 0.0;
  for (int64_t p = p0; p < p1; ++p) s += todbl_any(gu[p * ld + k]) * r1[p] - todbl_any(gv[p * ld + k]) * r2[p];
  partial[(int64_t)blockIdx.y * K + k] = s;
}
