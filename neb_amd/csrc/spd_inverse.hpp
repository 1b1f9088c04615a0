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

// Pivot block inversion: Pi = (A[k0:k0+nb, k0:k0+nb])^-1 by scalar sweeps.
// The 64 x 64 block lives in registers (thread t owns row t/4, columns
// 16*(t%4) .. +15); each column step exchanges only the pivot row and the
// pivot column through LDS.  Rows/columns >= nb are identity padding and are
// never swept.  info: 0 on success, else the 1-based global index of the
// first nonpositive / nonfinite pivot (first writer wins).
__global__ __launch_bounds__(256) void spd_pivot_kernel(const double* __restrict__ A, int64_t lda, int64_t k0,
                                                        int nb, double* __restrict__ Pi, int* __restrict__ info,
                                                        int reset) {
  __shared__ double rowbuf[2][SPD_NB];
  __shared__ double colbuf[2][SPD_NB];
  const int t = threadIdx.x;
  const int i = t >> 2, jg = (t & 3) * 16;
  if (reset && t == 0) *info = 0;  // first sweep clears the flag (no memset node)
  double v[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = jg + q;
    v[q] = (i < nb && j < nb) ? A[(k0 + i) * lda + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  int bad = 0;
  for (int c = 0; c < nb; ++c) {
    const int b = c & 1;  // alternate buffers: one barrier per column
    if (i == c) {
#pragma unroll
      for (int q = 0; q < 16; q += 2) *reinterpret_cast<f64x2*>(&rowbuf[b][jg + q]) = f64x2{v[q], v[q + 1]};
    }
    if ((c >> 4) == (t & 3)) {
      const int qc = c & 15;
      double cv = v[0];
#pragma unroll
      for (int q = 1; q < 16; ++q) cv = (q == qc) ? v[q] : cv;
      colbuf[b][i] = cv;
    }
    __syncthreads();
    double r[16];
#pragma unroll
    for (int q = 0; q < 16; q += 2) {
      const f64x2 rv = *reinterpret_cast<const f64x2*>(&rowbuf[b][jg + q]);
      r[q] = rv[0];
      r[q + 1] = rv[1];
    }
    double p = rowbuf[b][c];
    const double cc = colbuf[b][i];
    const bool okp = (p > 0.0) && isfinite(p);
    if (!okp && bad == 0) bad = (int)(k0 + c + 1);
    p = okp ? p : 1.0;  // keep the arithmetic finite; the caller raises on info
    const double ip = 1.0 / p;
    const double ci = cc * ip;
    const bool prow = i == c;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const bool pcol = (jg + q) == c;
      const double gen = pcol ? ci : v[q] - ci * r[q];
      const double rowv = pcol ? -ip : r[q] * ip;
      v[q] = prow ? rowv : gen;
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int j = jg + q;
    Pi[i * SPD_NB + j] = (i < nb && j < nb) ? -v[q] : 0.0;  // sweeps gave -P^-1
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
