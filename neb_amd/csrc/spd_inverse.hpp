// FP64 inverse of a symmetric positive definite K x K matrix by blocked
// symmetric Gauss-Jordan sweeps, with a positive-pivot test that is the
// positive-definiteness gate (every pivot of the sweep is a diagonal entry of
// a Schur complement, and all are > 0 exactly when the matrix is SPD).
//
// Replaces the FP64 Cholesky + triangular inverse of the spectral objective
// (velocity_trace_jet2.py:109-112 `cholesky` / `solve`, velocity/core.py:
// 52-56) for K up to a few thousand: the K^3 work is tiny, so the latency of
// the factorization's sequential steps is what matters; here there are
// 2 launches per 64 columns and each launch is fully parallel except the
// 64 x 64 pivot inversion, which one workgroup does in LDS.
//
// Sweep of pivot block k (P = A_kk, Pi = P^-1):
//   A_kk <- -Pi,  A_ik <- A_ik Pi,  A_kj <- Pi A_kj,  A_ij <- A_ij - A_ik Pi A_kj
// Sweeping every block yields -A^-1.  The scalar version of the same sweep
// inverts the pivot block inside the workgroup.
#pragma once
#include "neb_common.hpp"

constexpr int SPD_NB = 64;

// Pivot block inversion: Pi = (A[k0:k0+nb, k0:k0+nb])^-1 by scalar sweeps,
// lazily scaled.  The sweep's rank-1 updates commute with scaling a row or a
// column, so a swept row/column is left as it is (instead of multiplied by
// 1/p) and the pivot entry is only negated (stored -p for the true -1/p);
// afterwards every entry (i, j) carries the pending factor d_i d_j with
// d_c = 1/p_c.  Per column a thread then does one FMA per entry it holds:
//   v_ij -= x_i r_j,  x_i = v_ic / p (x_c = 0),  r_j = v_cj (r_c = 0).
// Pivots p are entries no earlier step has scaled, so the positivity test is
// the one of the plain sweep.  The 64 x 64 block lives in registers (thread t
// owns row t/4, columns 16*(t%4) .. +15); each column step exchanges the
// pivot row and column through LDS (double-buffered: one barrier per column).
// Rows/columns >= nb are identity padding and are never swept.  info: 0 on
// success, else the 1-based global index of the first nonpositive /
// nonfinite pivot (first writer wins).
__global__ __launch_bounds__(256) void spd_pivot_kernel(const double* __restrict__ A, int64_t lda, int64_t k0,
                                                        int nb, double* __restrict__ Pi, int* __restrict__ info,
                                                        int reset) {
  __shared__ __align__(16) double rowbuf[2][SPD_NB];
  __shared__ double colbuf[2][SPD_NB];
  __shared__ double dvec[SPD_NB];
  const int t = threadIdx.x;
  const int i = t >> 2, jb = t & 3, jg = jb * 16;
  if (reset && t == 0) *info = 0;  // first sweep clears the flag (no memset node)
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = jg + q;
    v[q] = (i < nb && j < nb) ? A[(k0 + i) * lda + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  int bad = 0;
  for (int c0 = 0; c0 < nb; c0 += 16) {
    const bool own = (c0 >> 4) == jb;  // this thread holds columns c0 .. c0+15
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int c = c0 + u, b = u & 1;
      if (c < nb) {  // uniform over the workgroup
        if (i == c) {
#pragma unroll
          for (int q = 0; q < 16; q += 2)
            *reinterpret_cast<f64x2*>(&rowbuf[b][jg + q]) =
                f64x2{(own && q == u) ? 0.0 : v[q], (own && q + 1 == u) ? 0.0 : v[q + 1]};
        }
        if (own) colbuf[b][i] = v[u];
        __syncthreads();
        double r[16];
#pragma unroll
        for (int q = 0; q < 16; q += 2) {
          const f64x2 rv = *reinterpret_cast<const f64x2*>(&rowbuf[b][jg + q]);
          r[q] = rv[0];
          r[q + 1] = rv[1];
        }
        const double p = colbuf[b][c];
        const bool okp = (p > 0.0) && isfinite(p);
        if (!okp && bad == 0) bad = (int)(k0 + c + 1);
        const double ip = 1.0 / (okp ? p : 1.0);  // keep the arithmetic finite; the caller raises on info
        const double x = (i == c) ? 0.0 : colbuf[b][i] * ip;
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = fma(-x, r[q], v[q]);
        v[u] = (own && i == c) ? -v[u] : v[u];
        if (t == 0) dvec[c] = ip;
      }
    }
  }
  __syncthreads();
  const double di = i < nb ? dvec[i] : 0.0;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = jg + q;
    Pi[i * SPD_NB + j] = (i < nb && j < nb) ? -v[q] * di * dvec[j] : 0.0;  // sweeps gave -P^-1
  }
  if (t == 0 && bad != 0) atomicCAS(info, 0, bad);
}

// acc[a][b] += sum_q LT[q][ti+a] * R[q][tj+b]  (LT = left operand transposed)
__device__ inline void spd_mm(const double (*LT)[SPD_NB + 2], const double (*R)[SPD_NB + 2], int ti, int tj, int nb,
                              double (&acc)[4][4]) {
  for (int q = 0; q < nb; ++q) {
    const f64x2 l01 = *reinterpret_cast<const f64x2*>(&LT[q][ti]);
    const f64x2 l23 = *reinterpret_cast<const f64x2*>(&LT[q][ti + 2]);
    const f64x2 r01 = *reinterpret_cast<const f64x2*>(&R[q][tj]);
    const f64x2 r23 = *reinterpret_cast<const f64x2*>(&R[q][tj + 2]);
    const double l[4] = {l01[0], l01[1], l23[0], l23[1]};
    const double r[4] = {r01[0], r01[1], r23[0], r23[1]};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += l[a] * r[b];
  }
}

// One block sweep over the whole matrix: B = sweep_k(A) (ping-pong buffers),
// scaled by `out_scale` (-1 on the last sweep turns -A^-1 into A^-1).
// 64 x 64 output tile per workgroup, 4 x 4 outputs per thread; LDS operands
// are stored [q][row] so every inner-product step reads 16-byte vectors.
__global__ __launch_bounds__(256) void spd_update_kernel(const double* __restrict__ A, int64_t lda, int64_t K,
                                                         int64_t k0, int nb, const double* __restrict__ Pi,
                                                         double* __restrict__ B, int64_t ldb, double out_scale) {
  extern __shared__ __align__(16) double spd_smem[];
  double(*L0)[SPD_NB + 2] = reinterpret_cast<double(*)[SPD_NB + 2]>(spd_smem);
  double(*L1)[SPD_NB + 2] = reinterpret_cast<double(*)[SPD_NB + 2]>(spd_smem + SPD_NB * (SPD_NB + 2));
  const int64_t i0 = (int64_t)blockIdx.y * SPD_NB, j0 = (int64_t)blockIdx.x * SPD_NB;
  const int t = threadIdx.x;
  const int ti = (t / 16) * 4, tj = (t % 16) * 4;  // this thread's 4x4 outputs
  const bool ipiv = i0 == k0, jpiv = j0 == k0;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;

  if (ipiv && jpiv) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = -Pi[(ti + a) * SPD_NB + tj + b];
  } else if (jpiv) {
    // A_ik Pi : LT[q][r] = A[i0+r][k0+q], R = Pi
    for (int e = t; e < SPD_NB * SPD_NB; e += 256) {
      const int r = e / SPD_NB, c = e % SPD_NB;
      L0[c][r] = (i0 + r < K && c < nb) ? A[(i0 + r) * lda + k0 + c] : 0.0;
      L1[r][c] = Pi[e];
    }
    __syncthreads();
    spd_mm(L0, L1, ti, tj, nb, acc);
  } else if (ipiv) {
    // Pi A_kj : LT = Pi^T = Pi (symmetric), R[q][c] = A[k0+q][j0+c]
    for (int e = t; e < SPD_NB * SPD_NB; e += 256) {
      const int r = e / SPD_NB, c = e % SPD_NB;
      L0[r][c] = Pi[e];
      L1[r][c] = (r < nb && j0 + c < K) ? A[(k0 + r) * lda + j0 + c] : 0.0;
    }
    __syncthreads();
    spd_mm(L0, L1, ti, tj, nb, acc);
  } else {
    // A_ij - (A_ik Pi) A_kj
    for (int e = t; e < SPD_NB * SPD_NB; e += 256) {
      const int r = e / SPD_NB, c = e % SPD_NB;
      L0[c][r] = (i0 + r < K && c < nb) ? A[(i0 + r) * lda + k0 + c] : 0.0;
      L1[r][c] = Pi[e];
    }
    __syncthreads();
    double tt[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) tt[a][b] = 0.0;
    spd_mm(L0, L1, ti, tj, nb, tt);  // T[ti+a][tj+b] = (A_ik Pi)
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) L0[tj + b][ti + a] = tt[a][b];  // LT[q][r] = T[r][q]
    for (int e = t; e < SPD_NB * SPD_NB; e += 256) {
      const int r = e / SPD_NB, c = e % SPD_NB;
      L1[r][c] = (r < nb && j0 + c < K) ? A[(k0 + r) * lda + j0 + c] : 0.0;
    }
    __syncthreads();
    double upd[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) upd[a][b] = 0.0;
    spd_mm(L0, L1, ti, tj, nb, upd);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int64_t gi = i0 + ti + a, gj = j0 + tj + b;
        acc[a][b] = ((gi < K && gj < K) ? A[gi * lda + gj] : 0.0) - upd[a][b];
      }
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t gi = i0 + ti + a, gj = j0 + tj + b;
      if (gi < K && gj < K) B[gi * ldb + gj] = out_scale * acc[a][b];
    }
}

constexpr size_t spd_update_lds_bytes() { return 2 * sizeof(double) * SPD_NB * (SPD_NB + 2); }
