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

// 4 x 4 inverse of a symmetric positive definite P by cofactors (one
// division), and its leading principal minors m1..m4 (Sylvester: P is SPD
// exactly when all four are positive; the first nonpositive one names the
// first nonpositive pivot of the scalar sweep, pivot_k = m_k / m_(k-1)).
__device__ inline void spd_inv4(const double (&a)[4][4], double (&inv)[4][4], double (&mn)[4]) {
  const double s0 = a[0][0] * a[1][1] - a[1][0] * a[0][1], s1 = a[0][0] * a[1][2] - a[1][0] * a[0][2];
  const double s2 = a[0][0] * a[1][3] - a[1][0] * a[0][3], s3 = a[0][1] * a[1][2] - a[1][1] * a[0][2];
  const double s4 = a[0][1] * a[1][3] - a[1][1] * a[0][3], s5 = a[0][2] * a[1][3] - a[1][2] * a[0][3];
  const double c5 = a[2][2] * a[3][3] - a[3][2] * a[2][3], c4 = a[2][1] * a[3][3] - a[3][1] * a[2][3];
  const double c3 = a[2][1] * a[3][2] - a[3][1] * a[2][2], c2 = a[2][0] * a[3][3] - a[3][0] * a[2][3];
  const double c1 = a[2][0] * a[3][2] - a[3][0] * a[2][2], c0 = a[2][0] * a[3][1] - a[3][0] * a[2][1];
  const double det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
  mn[0] = a[0][0];
  mn[1] = s0;
  mn[2] = a[2][2] * s0 - a[2][1] * s1 + a[2][0] * s3;
  mn[3] = det;
  const double id = 1.0 / det;
  inv[0][0] = (a[1][1] * c5 - a[1][2] * c4 + a[1][3] * c3) * id;
  inv[0][1] = (-a[0][1] * c5 + a[0][2] * c4 - a[0][3] * c3) * id;
  inv[0][2] = (a[3][1] * s5 - a[3][2] * s4 + a[3][3] * s3) * id;
  inv[0][3] = (-a[2][1] * s5 + a[2][2] * s4 - a[2][3] * s3) * id;
  inv[1][0] = (-a[1][0] * c5 + a[1][2] * c2 - a[1][3] * c1) * id;
  inv[1][1] = (a[0][0] * c5 - a[0][2] * c2 + a[0][3] * c1) * id;
  inv[1][2] = (-a[3][0] * s5 + a[3][2] * s2 - a[3][3] * s1) * id;
  inv[1][3] = (a[2][0] * s5 - a[2][2] * s2 + a[2][3] * s1) * id;
  inv[2][0] = (a[1][0] * c4 - a[1][1] * c2 + a[1][3] * c0) * id;
  inv[2][1] = (-a[0][0] * c4 + a[0][1] * c2 - a[0][3] * c0) * id;
  inv[2][2] = (a[3][0] * s4 - a[3][1] * s2 + a[3][3] * s0) * id;
  inv[2][3] = (-a[2][0] * s4 + a[2][1] * s2 - a[2][3] * s0) * id;
  inv[3][0] = (-a[1][0] * c3 + a[1][1] * c1 - a[1][2] * c0) * id;
  inv[3][1] = (a[0][0] * c3 - a[0][1] * c1 + a[0][2] * c0) * id;
  inv[3][2] = (-a[3][0] * s3 + a[3][1] * s1 - a[3][2] * s0) * id;
  inv[3][3] = (a[2][0] * s3 - a[2][1] * s1 + a[2][2] * s0) * id;
}

// Pivot block inversion: Pi = (A[k0:k0+nb, k0:k0+nb])^-1 by lazily scaled
// sweeps over 4 x 4 pivot blocks.  Sweeping block b with pivot block P
// (current values) is the rank-4 update
//   v_ij -= x_i r_j,  x_i = v_(i,b) P^-1 (0 for rows of b),  r_j = v_(b,j)
//   (0 for columns of b),
// after which the diagonal block is negated; rows and columns of b are left
// as they are.  The update commutes with left-multiplying a row block or
// right-multiplying a column block by a 4 x 4 matrix, so the pending
// factors P_b^-1 of every swept row / column block are applied once at the
// end: -A^-1 = diag(P^-1) * v * diag(P^-1).  Each step's P is a block no
// earlier step has transformed, so its leading minors are those of the
// plain sweep (the positive-definiteness gate).  The 64 x 64 block lives in
// registers (thread t owns row t/4, columns 16*(t%4) .. +15); each step
// exchanges the pivot rows and columns through LDS (double-buffered: one
// barrier per 4 columns).  Rows/columns >= nb are identity padding.  info:
// 0 on success, else the 1-based global index of the first nonpositive /
// nonfinite pivot (first writer wins).
__global__ __launch_bounds__(256) void spd_pivot_kernel(const double* __restrict__ A, int64_t lda, int64_t k0,
                                                        int nb, double* __restrict__ Pi, int* __restrict__ info,
                                                        int reset) {
  __shared__ __align__(16) double rowbuf[2][4][SPD_NB];
  __shared__ __align__(16) double colbuf[2][SPD_NB][4];
  __shared__ __align__(16) double pivs[SPD_NB / 4][16];
  __shared__ __align__(16) double M[SPD_NB][SPD_NB + 2];
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
    for (int s = 0; s < 4; ++s) {
      const int c = c0 + 4 * s, u = 4 * s, b = s & 1;
      if (c < nb) {  // uniform over the workgroup
        const bool inblk = i >= c && i < c + 4;
        if (inblk) {
#pragma unroll
          for (int q = 0; q < 16; q += 2)
            *reinterpret_cast<f64x2*>(&rowbuf[b][i - c][jg + q]) =
                f64x2{(own && q >= u && q < u + 4) ? 0.0 : v[q], (own && q + 1 >= u && q + 1 < u + 4) ? 0.0 : v[q + 1]};
        }
        if (own) {
          *reinterpret_cast<f64x2*>(&colbuf[b][i][0]) = f64x2{v[u], v[u + 1]};
          *reinterpret_cast<f64x2*>(&colbuf[b][i][2]) = f64x2{v[u + 2], v[u + 3]};
        }
        __syncthreads();
        double P[4][4], Q[4][4], mn[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const f64x2 p01 = *reinterpret_cast<const f64x2*>(&colbuf[b][c + r][0]);
          const f64x2 p23 = *reinterpret_cast<const f64x2*>(&colbuf[b][c + r][2]);
          P[r][0] = p01[0]; P[r][1] = p01[1]; P[r][2] = p23[0]; P[r][3] = p23[1];
        }
        spd_inv4(P, Q, mn);
        bool ok = true;
        int first = 4;
#pragma unroll
        for (int r = 3; r >= 0; --r)
          if (!(mn[r] > 0.0) || !isfinite(mn[r])) { ok = false; first = r; }
        if (!ok) {
          if (bad == 0) bad = (int)(k0 + c + first + 1);
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) Q[r][q] = r == q ? 1.0 : 0.0;  // keep the arithmetic finite
        }
        const f64x2 a01 = *reinterpret_cast<const f64x2*>(&colbuf[b][i][0]);
        const f64x2 a23 = *reinterpret_cast<const f64x2*>(&colbuf[b][i][2]);
        const double ai[4] = {a01[0], a01[1], a23[0], a23[1]};
        double x[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const double xm = ai[0] * Q[0][m] + ai[1] * Q[1][m] + ai[2] * Q[2][m] + ai[3] * Q[3][m];
          x[m] = inblk ? 0.0 : xm;
        }
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
          for (int q = 0; q < 16; q += 2) {
            const f64x2 rv = *reinterpret_cast<const f64x2*>(&rowbuf[b][m][jg + q]);
            v[q] = fma(-x[m], rv[0], v[q]);
            v[q + 1] = fma(-x[m], rv[1], v[q + 1]);
          }
        }
        if (own && inblk) {
#pragma unroll
          for (int k = 0; k < 4; ++k) v[u + k] = -v[u + k];
        }
        if (t == 0) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) pivs[c >> 2][4 * r + q] = Q[r][q];
        }
      }
    }
  }
  // pending factors: T = diag(P^-1) v (rows) through LDS, then T diag(P^-1)
  // (columns) in registers; Pi = -T
#pragma unroll
  for (int q = 0; q < 16; q += 2) *reinterpret_cast<f64x2*>(&M[i][jg + q]) = f64x2{v[q], v[q + 1]};
  __syncthreads();
  const int nbk = (nb + 3) >> 2;
  const int I = i >> 2, ir = i & 3;
  double tr[16];
  if (I < nbk) {
    double li[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) li[a] = pivs[I][4 * ir + a];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      double acc = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) acc = fma(li[a], M[4 * I + a][jg + q], acc);
      tr[q] = acc;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) tr[q] = 0.0;
  }
#pragma unroll
  for (int J = 0; J < 4; ++J) {
    const int Jg = (jg >> 2) + J;  // global column block
    double o[4] = {0.0, 0.0, 0.0, 0.0};
    if (Jg < nbk) {
#pragma unroll
      for (int bq = 0; bq < 4; ++bq)
#pragma unroll
        for (int cq = 0; cq < 4; ++cq) o[cq] = fma(tr[4 * J + bq], pivs[Jg][4 * bq + cq], o[cq]);
    }
#pragma unroll
    for (int cq = 0; cq < 4; ++cq) {
      const int j = jg + 4 * J + cq;
      Pi[i * SPD_NB + j] = (i < nb && j < nb) ? -o[cq] : 0.0;
    }
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
