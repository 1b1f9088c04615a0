// FP64 inverse of a symmetric positive definite K x K matrix by blocked
// symmetric Gauss-Jordan sweeps, with a positive-pivot test that is the
// positive-definiteness gate (every pivot of the sweep is a diagonal entry of
// a Schur complement, and all are > 0 exactly when the matrix is SPD).
//
// Replaces the FP64 Cholesky + triangular inverse of the spectral objective
// (velocity_trace_jet2.py:109-112 `cholesky` / `solve`, velocity/core.py:
// 52-56) for K up to a few thousand: the K^3 work is tiny, so the latency of
// the sequential steps is what matters.  There is one launch per 64 columns
// (plus one for the first pivot): launch k applies sweep k to every 64 x 64
// tile, and the workgroup that produces tile (k+1, k+1) inverts it in the
// same launch, so the next launch finds its pivot inverse ready.  Products
// run on v_mfma_f64_16x16x4f64 from LDS tiles; the pivot inversion keeps its
// block in MFMA registers and applies its rank-4 updates as MFMAs.
//
// Sweep of pivot block k (P = A_kk, Pi = P^-1):
//   A_kk <- -Pi,  A_ik <- A_ik Pi,  A_kj <- Pi A_kj,  A_ij <- A_ij - A_ik Pi A_kj
// Sweeping every block yields -A^-1.  The matrix is swept padded to
// Kp = 64 ceil(K / 64) with an identity block beyond K (its sweep leaves the
// identity in place and the K x K part untouched), so every tile is full.
// The sweeps keep the matrix symmetric, so A_kj is read as A_jk^T: every
// operand is a tile of block column k, stored [row][k] in LDS.
//
// The same sweeps also yield T = U^-1, the inverse of the upper Cholesky
// factor (A = U^T U, T^T A T = I, T upper with a positive diagonal), which
// the coefficient peel needs (velocity/core.py:53-56).  Before sweep k, block
// column k above the diagonal holds A_<k^-1 A_<k,k and the diagonal block
// holds the Schur complement S_k; with U_<k^T U_<k,k = A_<k,k and
// U_kk^T U_kk = S_k, block column k of T is
//   T_<k,k = -(A_<k^-1 A_<k,k) T_kk,   T_kk = U_kk^-1,
// and the pivot workgroup forms T_kk beside the inverse of S_k from the same
// 4-column sweep (spd_pivot_regs).
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


constexpr int SPD_LDP = SPD_NB + 4;         // LDS row stride (doubles): conflict-free 16-byte operand reads
constexpr int SPD_TILE = SPD_NB * SPD_LDP;  // doubles per LDS tile
// pivot inversion scratch: pivot rows [2][4][64], pivot columns [2][64][4],
// the 16 block inverses [16][16] and the final 64 x 64 tile
constexpr int SPD_PIVOT_LDS = 2 * 4 * SPD_NB + 2 * SPD_NB * 4 + 16 * 16 + SPD_TILE;

// Register tile in MFMA layout: wave w owns the 32 x 32 quadrant at rows
// 32 (w / 2), columns 32 (w % 2) as 2 x 2 tiles of 16 x 16; element e of
// tile (tr, tc) in lane l is quadrant row 16 tr + l / 16 + 4 e, column
// 16 tc + l % 16 (the v_mfma_f64_16x16x4f64 result layout).
__device__ inline int spd_row(int R0, int tr, int lane, int e) { return R0 + 16 * tr + (lane >> 4) + 4 * e; }
__device__ inline int spd_col(int C0, int tc, int lane) { return C0 + 16 * tc + (lane & 15); }

// Thread t's share of a 64 x 64 tile (rows i0.., columns j0..): pair u holds
// row (t + 256 u) / 32, columns 2 ((t + 256 u) % 32) + {0, 1}.  bounded: the
// source is the caller's K x K matrix, read as identity beyond K; otherwise a
// padded buffer with full tiles.
struct SpdPart {
  f64x2 v[8];
};

__device__ inline SpdPart spd_fetch(const double* __restrict__ src, int64_t ld, int64_t K, bool bounded, int64_t i0,
                                    int64_t j0, int t) {
  SpdPart p;
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int q = t + 256 * u;
    const int64_t gi = i0 + (q >> 5), gj = j0 + 2 * (q & 31);
    if (bounded) {
      p.v[u][0] = (gi < K && gj < K) ? src[gi * ld + gj] : (gi == gj ? 1.0 : 0.0);
      p.v[u][1] = (gi < K && gj + 1 < K) ? src[gi * ld + gj + 1] : (gi == gj + 1 ? 1.0 : 0.0);
    } else {
      p.v[u] = *reinterpret_cast<const f64x2*>(src + gi * ld + gj);
    }
  }
  return p;
}

__device__ inline void spd_put(double* L, const SpdPart& p, int t) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const int q = t + 256 * u;
    *reinterpret_cast<f64x2*>(L + (q >> 5) * SPD_LDP + 2 * (q & 31)) = p.v[u];
  }
}

// acc[tr][tc] += L[quadrant rows][0:64] * RT[quadrant columns][0:64]^T with
// L and RT stored [row][k] in LDS.  MFMA pair p takes k = 8p + 2g + e in its
// e-th instruction (lane group g = lane / 16) for both operands, so each lane
// reads 16 bytes per operand per pair.
__device__ inline void spd_mma(const double* L, const double* RT, int R0, int C0, int lane, f64x4 (&acc)[2][2]) {
  const double* la = L + (R0 + (lane & 15)) * SPD_LDP + 2 * (lane >> 4);
  const double* lb = RT + (C0 + (lane & 15)) * SPD_LDP + 2 * (lane >> 4);
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    const f64x2 a0 = *reinterpret_cast<const f64x2*>(la + 8 * p);
    const f64x2 a1 = *reinterpret_cast<const f64x2*>(la + 16 * SPD_LDP + 8 * p);
    const f64x2 b0 = *reinterpret_cast<const f64x2*>(lb + 8 * p);
    const f64x2 b1 = *reinterpret_cast<const f64x2*>(lb + 16 * SPD_LDP + 8 * p);
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[e], b0[e], acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[e], b1[e], acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[e], b0[e], acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[e], b1[e], acc[1][1], 0, 0, 0);
    }
  }
}

__device__ inline void spd_acc_store(double* M, const f64x4 (&v)[2][2], int R0, int C0, int lane) {
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
      for (int e = 0; e < 4; ++e) M[spd_row(R0, tr, lane, e) * SPD_LDP + spd_col(C0, tc, lane)] = v[tr][tc][e];
}

// Inversion of the 64 x 64 pivot block held in MFMA layout (v) by lazily
// scaled sweeps over 4 x 4 pivot blocks.  Sweeping block b (columns
// c = 4b .. c+3) with pivot block P (current values) is the rank-4 update
//   v_ij -= x_i r_j,  x_i = v_(i,b) P^-1 (0 for rows of b),
//   r_j = v_(b,j) (0 for columns of b),
// one MFMA per 16 x 16 tile with the four pivot columns as its k, after
// which the diagonal block is negated; rows and columns of b are left as
// they are.  The update commutes with left-multiplying a row block or
// right-multiplying a column block by a 4 x 4 matrix, so the pending factors
// P_b^-1 of every swept row / column block are applied once at the end:
// -A^-1 = diag(P^-1) v diag(P^-1).  Each step's P is a block no earlier step
// has transformed, so its leading minors are those of the plain sweep (the
// positive-definiteness gate).  Pivot rows and columns pass through LDS,
// double-buffered: one barrier per 4 columns.  lds: SPD_PIVOT_LDS doubles,
// not in use by the caller.  Writes Pi = P^-1 (64 x 64, row stride 64).
// info: the 1-based index k0 + i + 1 of the first nonpositive / nonfinite
// pivot (first writer wins).
//
// TkT (optional): T_kk^T for T_kk = U^-1, U^T U = the pivot block.  By the
// identity of the header comment at 4-column granularity, column block b of
// T_kk is [-(S_<b^-1 S_<b,b) T_bb ; T_bb] with T_bb the inverse upper
// Cholesky factor of the 4 x 4 pivot P_b; with the pending row factors the
// swept block (a, b) is P_a^-1 v_(a,b), so T_(a,b) = -P_a^-1 v_(a,b) T_bb.
// Each step leaves its column block (the lazy v_(i,b), P_b on the diagonal)
// in M; after the last step the 16 T_bb are factored side by side and the
// blocks formed, off the sweep's dependency chain.
__device__ inline void spd_pivot_regs(f64x4 (&v)[2][2], double* lds, int w, int lane, int t, int64_t k0,
                                      double* __restrict__ Pi, int* __restrict__ info, double* __restrict__ TkT) {
  double* rowbuf = lds;
  double* colbuf = lds + 2 * 4 * SPD_NB;
  double* pivs = colbuf + 2 * SPD_NB * 4;
  double* M = pivs + 16 * 16;
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  const int g = lane >> 4, r16 = lane & 15;
  int bad = 0;
  for (int b4 = 0; b4 < 4; ++b4) {
    const int hq = b4 >> 1;  // quadrant half holding these 16 rows / columns
    const bool tq = b4 & 1;  // 16 x 16 tile inside that half
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int b = 4 * b4 + s, c = 4 * b;
      double* rb = rowbuf + (b & 1) * 4 * SPD_NB;
      double* cb = colbuf + (b & 1) * SPD_NB * 4;
      if ((w >> 1) == hq) {  // pivot row c + g is element s of tile row tq
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) rb[g * SPD_NB + spd_col(C0, tc, lane)] = tq ? v[1][tc][s] : v[0][tc][s];
      }
      if ((w & 1) == hq && (r16 >> 2) == s) {  // pivot column c + r16 % 4 is lane column r16 of tile column tq
#pragma unroll
        for (int tr = 0; tr < 2; ++tr)
#pragma unroll
          for (int e = 0; e < 4; ++e) cb[spd_row(R0, tr, lane, e) * 4 + (r16 & 3)] = tq ? v[tr][1][e] : v[tr][0][e];
      }
      __syncthreads();
      double P[4][4], Q[4][4], mn[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const f64x2 p01 = *reinterpret_cast<const f64x2*>(&cb[(c + r) * 4]);
        const f64x2 p23 = *reinterpret_cast<const f64x2*>(&cb[(c + r) * 4 + 2]);
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
      // k of the MFMA = lane group g: x_i[g] = v_(i,b) P^-1[:, g], -r_g(j).
      // Lane-dependent choices are exact arithmetic masks (x 1.0 / x 0.0 of
      // finite values), not selects: the compiler turns those into divergent
      // branches with an LDS wait inside.
      const double e0 = g == 0 ? 1.0 : 0.0, e1 = g == 1 ? 1.0 : 0.0, e2 = g == 2 ? 1.0 : 0.0, e3 = g == 3 ? 1.0 : 0.0;
      double qc[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) qc[a] = Q[a][0] * e0 + Q[a][1] * e1 + Q[a][2] * e2 + Q[a][3] * e3;
      double x[2], y[2];
#pragma unroll
      for (int tr = 0; tr < 2; ++tr) {
        const int i = R0 + 16 * tr + r16;
        const f64x2 a01 = *reinterpret_cast<const f64x2*>(&cb[i * 4]);
        const f64x2 a23 = *reinterpret_cast<const f64x2*>(&cb[i * 4 + 2]);
        if (TkT && g == 0 && (w & 1) == 0) {  // waves 0 and 2 cover the 64 rows: snapshot of column block b
          *reinterpret_cast<f64x2*>(&M[i * SPD_LDP + c]) = a01;
          *reinterpret_cast<f64x2*>(&M[i * SPD_LDP + c + 2]) = a23;
        }
        const double xv = a01[0] * qc[0] + a01[1] * qc[1] + a23[0] * qc[2] + a23[1] * qc[3];
        x[tr] = xv * ((i >> 2) == b ? 0.0 : 1.0);
      }
#pragma unroll
      for (int tc = 0; tc < 2; ++tc) {
        const int j = spd_col(C0, tc, lane);
        y[tc] = -rb[g * SPD_NB + j] * ((j >> 2) == b ? 0.0 : 1.0);
      }
#pragma unroll
      for (int tr = 0; tr < 2; ++tr)
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) v[tr][tc] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[tr], y[tc], v[tr][tc], 0, 0, 0);
      if ((w >> 1) == hq && (w & 1) == hq && (r16 >> 2) == s) {  // negate the diagonal pivot block
        if (tq) v[1][1][s] = -v[1][1][s];
        else v[0][0][s] = -v[0][0][s];
      }
      if (w == 0 && r16 == 0) {  // lane 16 g stores column g of P^-1
#pragma unroll
        for (int r = 0; r < 4; ++r) pivs[b * 16 + 4 * r + g] = qc[r];
      }
    }
  }
  if (TkT) {
    __syncthreads();  // every column snapshot is in M
    double* tb = rowbuf;  // T_bb of the 16 pivots, [b][4][4]
    if (t < 16) {         // U^T U = P_b (upper U), T_bb = U^-1
      const double* pb = M + 4 * t * SPD_LDP + 4 * t;
      const double p00 = pb[0], p01 = pb[1], p02 = pb[2], p03 = pb[3];
      const double p11 = pb[SPD_LDP + 1], p12 = pb[SPD_LDP + 2], p13 = pb[SPD_LDP + 3];
      const double p22 = pb[2 * SPD_LDP + 2], p23 = pb[2 * SPD_LDP + 3], p33 = pb[3 * SPD_LDP + 3];
      const double u00 = sqrt(p00), i0 = 1.0 / u00;
      const double u01 = p01 * i0, u02 = p02 * i0, u03 = p03 * i0;
      const double u11 = sqrt(p11 - u01 * u01), i1 = 1.0 / u11;
      const double u12 = (p12 - u01 * u02) * i1, u13 = (p13 - u01 * u03) * i1;
      const double u22 = sqrt(p22 - u02 * u02 - u12 * u12), i2 = 1.0 / u22;
      const double u23 = (p23 - u02 * u03 - u12 * u13) * i2;
      const double i3 = 1.0 / sqrt(p33 - u03 * u03 - u13 * u13 - u23 * u23);
      const double t23 = -u23 * i3 * i2;
      const double t12 = -u12 * i2 * i1, t13 = -(u12 * t23 + u13 * i3) * i1;
      const double t01 = -u01 * i1 * i0, t02 = -(u01 * t12 + u02 * i2) * i0;
      const double t03 = -(u01 * t13 + u02 * t23 + u03 * i3) * i0;
      double* o = tb + 16 * t;
      o[0] = i0;  o[1] = t01; o[2] = t02;  o[3] = t03;
      o[4] = 0.0; o[5] = i1;  o[6] = t12;  o[7] = t13;
      o[8] = 0.0; o[9] = 0.0; o[10] = i2;  o[11] = t23;
      o[12] = 0.0; o[13] = 0.0; o[14] = 0.0; o[15] = i3;
    }
    __syncthreads();
    // thread t: row i = t / 4 of T_kk, columns jg .. jg+15 (column blocks jb .. jb+3)
    const int i = t >> 2, a = i >> 2, ir = i & 3, jg = (t & 3) * 16;
    double pa[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) pa[q] = pivs[a * 16 + 4 * ir + q];  // row ir of P_a^-1
#pragma unroll
    for (int J = 0; J < 4; ++J) {
      const int b = (jg >> 2) + J;
      double tv[4];
      if (b < a) {
#pragma unroll
        for (int q = 0; q < 4; ++q) tv[q] = 0.0;
      } else if (b == a) {
#pragma unroll
        for (int q = 0; q < 4; ++q) tv[q] = tb[16 * b + 4 * ir + q];
      } else {  // -(P_a^-1 v_(a,b) T_bb) row ir
        double hv[4] = {0.0, 0.0, 0.0, 0.0};  // row ir of P_a^-1 v_(a,b)
#pragma unroll
        for (int s2 = 0; s2 < 4; ++s2)
#pragma unroll
          for (int u = 0; u < 4; ++u) hv[u] = fma(pa[s2], M[(4 * a + s2) * SPD_LDP + 4 * b + u], hv[u]);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          double acc = 0.0;
#pragma unroll
          for (int u = 0; u < 4; ++u) acc = fma(hv[u], tb[16 * b + 4 * u + q], acc);
          tv[q] = -acc;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) TkT[(4 * b + q) * SPD_NB + i] = tv[q];
    }
    __syncthreads();  // M is reused below
  }
  // pending factors: T = diag(P^-1) v (rows) through LDS, then T diag(P^-1)
  // (columns) in registers; Pi = -T
  spd_acc_store(M, v, R0, C0, lane);
  __syncthreads();
  const int i = t >> 2, jg = (t & 3) * 16;
  const int I = i >> 2, ir = i & 3;
  double li[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) li[a] = pivs[I * 16 + 4 * ir + a];
  double tr_[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) acc = fma(li[a], M[(4 * I + a) * SPD_LDP + jg + q], acc);
    tr_[q] = acc;
  }
#pragma unroll
  for (int J = 0; J < 4; ++J) {
    const int Jg = (jg >> 2) + J;  // column block
    double o[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int bq = 0; bq < 4; ++bq)
#pragma unroll
      for (int cq = 0; cq < 4; ++cq) o[cq] = fma(tr_[4 * J + bq], pivs[Jg * 16 + 4 * bq + cq], o[cq]);
    *reinterpret_cast<f64x2*>(&Pi[i * SPD_NB + jg + 4 * J]) = f64x2{-o[0], -o[1]};
    *reinterpret_cast<f64x2*>(&Pi[i * SPD_NB + jg + 4 * J + 2]) = f64x2{-o[2], -o[3]};
  }
  if (t == 0 && bad != 0) atomicCAS(info, 0, bad);
}

// First pivot: Pi = (A[0:64, 0:64], identity beyond K)^-1; clears info.
// TkT (optional): T_00^T (spd_pivot_regs).
__global__ __launch_bounds__(256) void spd_pivot0_kernel(const double* __restrict__ A, int64_t lda, int64_t K,
                                                         double* __restrict__ Pi, int* __restrict__ info,
                                                         double* __restrict__ TkT) {
  extern __shared__ __align__(16) double spd_smem[];
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  if (t == 0) *info = 0;  // the first launch of an inversion clears the flag (no memset node)
  spd_put(spd_smem, spd_fetch(A, lda, K, true, 0, 0, t), t);
  __syncthreads();
  f64x4 v[2][2];
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        v[tr][tc][e] = spd_smem[spd_row(R0, tr, lane, e) * SPD_LDP + spd_col(C0, tc, lane)];
  __syncthreads();
  spd_pivot_regs(v, spd_smem, w, lane, t, 0, Pi, info, TkT);
}

// Sweep k on tile pair (ti, tj), ti <= tj (blockIdx.x counts the upper
// triangle of the nt x nt tile grid row by row) of the padded matrix,
// src -> dst: the sweeps keep the matrix symmetric, so the workgroup stores
// its tile and, through an LDS transpose, the mirrored tile (tj, ti).  On
// the last sweep dst is X: tiles are negated (-A^-1 -> A^-1) and clipped to
// K x K.  The workgroup of tile (k+1, k+1) inverts its new value as the next
// pivot (Pi_next) instead of storing it: sweep k+1 replaces that tile by
// -Pi_next without reading it.
//
// With Tout (K x K, row stride ldt) the launch also writes block column k of
// T = U^-1 (header comment): tile (i, k), i < k, = -(A_ik) T_kk from the
// tile the workgroup reads anyway, tile (k, k) = T_kk, and the workgroups of
// row k write the zero tiles (j, k), j > k; T_kk^T comes in TkT, and the
// pivot workgroup leaves T_k+1,k+1^T in TkT_next.
__global__ __launch_bounds__(256) void spd_step_kernel(const double* __restrict__ src, int64_t lds, int bounded,
                                                       int64_t K, int k, int nt, const double* __restrict__ Pi,
                                                       double* __restrict__ Pi_next, double* __restrict__ dst,
                                                       int64_t ldd, int last, int* __restrict__ info,
                                                       const double* __restrict__ TkT, double* __restrict__ TkT_next,
                                                       double* __restrict__ Tout, int64_t ldt) {
  // two LDS tiles (70 KB): a workgroup fits beside a 64 KB cross-GEMM
  // workgroup on a CU, so the sweeps share the GPU with the output-layer
  // products queued on another stream
  extern __shared__ __align__(16) double spd_smem[];
  double* L0 = spd_smem;
  double* L1 = L0 + SPD_TILE;
  const int t = threadIdx.x, lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int R0 = 32 * (w >> 1), C0 = 32 * (w & 1);
  int ti = 0, tj;
  {
    int u = blockIdx.x, rowlen = nt;
    while (u >= rowlen) { u -= rowlen; ++ti; --rowlen; }
    tj = ti + u;
  }
  const int64_t i0 = (int64_t)ti * SPD_NB, j0 = (int64_t)tj * SPD_NB, k0 = (int64_t)k * SPD_NB;
  const bool bnd = bounded != 0;
  f64x4 acc[2][2];
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc) acc[tr][tc] = f64x4{0.0, 0.0, 0.0, 0.0};
  if (ti == k && tj == k) {  // A_kk <- -Pi
#pragma unroll
    for (int tr = 0; tr < 2; ++tr)
#pragma unroll
      for (int tc = 0; tc < 2; ++tc)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[tr][tc][e] = -Pi[spd_row(R0, tr, lane, e) * SPD_NB + spd_col(C0, tc, lane)];
    if (Tout) {  // T_kk (row r, column c) = TkT[c][r]
#pragma unroll 4
      for (int u = 0; u < 16; ++u) {
        const int q = t + 256 * u, r = q >> 6, c = q & 63;
        if (k0 + r < K && k0 + c < K) Tout[(k0 + r) * ldt + k0 + c] = TkT[c * SPD_NB + r];
      }
    }
  } else if (tj == k) {  // A_ik Pi  (RT = Pi^T = Pi)
    const SpdPart pa = spd_fetch(src, lds, K, bnd, i0, k0, t);
    const SpdPart pp = spd_fetch(Pi, SPD_NB, SPD_NB, false, 0, 0, t);
    SpdPart pt;
    if (Tout) pt = spd_fetch(TkT, SPD_NB, SPD_NB, false, 0, 0, t);
    spd_put(L0, pa, t);
    spd_put(L1, pp, t);
    __syncthreads();
    spd_mma(L0, L1, R0, C0, lane, acc);
    if (Tout) {  // T_ik = -(A_ik) T_kk  (RT = T_kk^T)
      f64x4 tq[2][2];
#pragma unroll
      for (int tr = 0; tr < 2; ++tr)
#pragma unroll
        for (int tc = 0; tc < 2; ++tc) tq[tr][tc] = f64x4{0.0, 0.0, 0.0, 0.0};
      __syncthreads();  // every wave is done reading L1
      spd_put(L1, pt, t);
      __syncthreads();
      spd_mma(L0, L1, R0, C0, lane, tq);
#pragma unroll
      for (int tr = 0; tr < 2; ++tr)
#pragma unroll
        for (int tc = 0; tc < 2; ++tc)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int64_t gi = i0 + spd_row(R0, tr, lane, e), gj = k0 + spd_col(C0, tc, lane);
            if (gi < K && gj < K) Tout[gi * ldt + gj] = -tq[tr][tc][e];
          }
    }
  } else if (ti == k) {  // Pi A_kj = Pi A_jk^T
    if (Tout) {  // T is upper triangular: zero tile (j, k)
#pragma unroll 4
      for (int u = 0; u < 16; ++u) {
        const int q = t + 256 * u, r = q >> 6, c = q & 63;
        if (j0 + r < K && k0 + c < K) Tout[(j0 + r) * ldt + k0 + c] = 0.0;
      }
    }
    const SpdPart pp = spd_fetch(Pi, SPD_NB, SPD_NB, false, 0, 0, t);
    const SpdPart pb = spd_fetch(src, lds, K, bnd, j0, k0, t);
    spd_put(L0, pp, t);
    spd_put(L1, pb, t);
    __syncthreads();
    spd_mma(L0, L1, R0, C0, lane, acc);
  } else {  // A_ij - (A_ik Pi) A_jk^T
    const SpdPart pa = spd_fetch(src, lds, K, bnd, i0, k0, t);
    const SpdPart pp = spd_fetch(Pi, SPD_NB, SPD_NB, false, 0, 0, t);
    const SpdPart pb = spd_fetch(src, lds, K, bnd, j0, k0, t);
    f64x4 aij[2][2];
#pragma unroll
    for (int tr = 0; tr < 2; ++tr)
#pragma unroll
      for (int tc = 0; tc < 2; ++tc)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int64_t gi = i0 + spd_row(R0, tr, lane, e), gj = j0 + spd_col(C0, tc, lane);
          aij[tr][tc][e] = !bnd ? src[gi * lds + gj] : (gi < K && gj < K) ? src[gi * lds + gj] : (gi == gj ? 1.0 : 0.0);
        }
    spd_put(L0, pa, t);
    spd_put(L1, pp, t);
    __syncthreads();
    f64x4 tt[2][2];
#pragma unroll
    for (int tr = 0; tr < 2; ++tr)
#pragma unroll
      for (int tc = 0; tc < 2; ++tc) tt[tr][tc] = f64x4{0.0, 0.0, 0.0, 0.0};
    spd_mma(L0, L1, R0, C0, lane, tt);  // T = A_ik Pi
    __syncthreads();                     // every wave is done reading L0 and L1
    spd_acc_store(L0, tt, R0, C0, lane);
    spd_put(L1, pb, t);                  // A_jk (held in registers) replaces Pi
    __syncthreads();
    spd_mma(L0, L1, R0, C0, lane, acc);  // T A_jk^T
#pragma unroll
    for (int tr = 0; tr < 2; ++tr)
#pragma unroll
      for (int tc = 0; tc < 2; ++tc) acc[tr][tc] = aij[tr][tc] - acc[tr][tc];
    if (ti == k + 1 && tj == k + 1) {
      __syncthreads();  // the pivot inversion reuses the LDS tiles
      spd_pivot_regs(acc, spd_smem, w, lane, t, (int64_t)(k + 1) * SPD_NB, Pi_next, info, Tout ? TkT_next : nullptr);
      return;
    }
  }
  const double sgn = last ? -1.0 : 1.0;
#pragma unroll
  for (int tr = 0; tr < 2; ++tr)
#pragma unroll
    for (int tc = 0; tc < 2; ++tc)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t gi = i0 + spd_row(R0, tr, lane, e), gj = j0 + spd_col(C0, tc, lane);
        if (!last || (gi < K && gj < K)) dst[gi * ldd + gj] = sgn * acc[tr][tc][e];
      }
  if (ti != tj) {  // the mirrored tile (tj, ti), row-contiguous through LDS
    __syncthreads();  // every wave is done with the operand tiles
    spd_acc_store(L0, acc, R0, C0, lane);
    __syncthreads();
#pragma unroll 4
    for (int u = 0; u < 16; ++u) {
      const int q = t + 256 * u, r = q >> 6, c = q & 63;  // row r of the transpose = column r of the tile
      const int64_t gi = j0 + r, gj = i0 + c;
      if (!last || (gi < K && gj < K)) dst[gi * ldd + gj] = sgn * L0[c * SPD_LDP + r];
    }
  }
}

constexpr size_t spd_step_lds_bytes() { return 2 * sizeof(double) * SPD_TILE; }
static_assert(SPD_PIVOT_LDS <= 2 * SPD_TILE, "the pivot inversion reuses the sweep's LDS tiles");
constexpr size_t spd_pivot0_lds_bytes() { return sizeof(double) * SPD_PIVOT_LDS; }
