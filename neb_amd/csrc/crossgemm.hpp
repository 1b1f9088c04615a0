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
// CR_BF16X3 also takes operands already split into hi / lo bf16 planes
// (SplitBF16), which skips the split in the loader.
// MFMA accumulators are flushed into a second fp32 (fp64 for CR_F64)
// register set every FLUSH rows, which bounds the length of any fp32
// rounding chain; partials are then combined in FP64.
#pragma once
#include <type_traits>

#include "neb_common.hpp"

enum { CR_F32 = 0, CR_BF16 = 1, CR_BF16X3 = 2, CR_F64 = 3 };
constexpr int CR_MAXSEG = 8;
constexpr int CR_TM = 128, CR_TN = 128;

struct CrossSeg {
  const void* x;
  const void* y;
  int64_t ldx, ldy, rows;
  double sign;
  const void* xlo;  // lo planes of pre-split operands (SplitBF16), else null
  const void* ylo;
};

// Operand tag: an fp32 matrix given as two bf16 planes hi = bf16(x) and
// lo = bf16(x - hi) with the same leading dimension (CR_BF16X3 only).
struct SplitBF16 {};

struct CrossArgs {
  CrossSeg seg[CR_MAXSEG];
  int nseg;
  int64_t m, n;
  const double* roww;  // optional per-row weights (single segment only)
  int symmetric;       // upper tiles only (X == Y)
  int bdiag;           // > 0: block-diagonal tiles only, bdiag x bdiag tiles per block (m == n)
  int tiles_m, tiles_n, ntiles, splits;
  double* partial;     // [splits][ntiles][TM][TN]
};

// Output tile of work item `tile`: every tile, the upper triangle
// (symmetric), or the tiles of the diagonal blocks (bdiag tiles a side).
__host__ __device__ inline void cross_tile_ij(const CrossArgs& a, int tile, int& ti, int& tj) {
  if (a.symmetric) {
    int t = tile, rowlen = a.tiles_n;
    ti = 0;
    while (t >= rowlen) { t -= rowlen; ++ti; --rowlen; }
    tj = ti + t;
  } else if (a.bdiag > 0) {
    const int per = a.bdiag * a.bdiag, blk = tile / per, r = tile % per;
    ti = blk * a.bdiag + r / a.bdiag;
    tj = blk * a.bdiag + r % a.bdiag;
  } else {
    ti = tile / a.tiles_n;
    tj = tile % a.tiles_n;
  }
}

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
  static constexpr int KI = 32, KE = 8, RK = 32, PAD = 0, NPROD = 1, FLUSH = 1 << 30;
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

// Transposing tile loader: rows [r0, r0+RK) x cols [c0, c0+TM) of a
// row-major source -> LDS hi[col][row] (and lo[col][row] for the split
// mode).  Thread t owns column t % TM and RK/2 consecutive rows, so each
// global load instruction reads TM consecutive elements of one row.  fetch()
// fills registers (issued one stage ahead of the MFMAs); commit() converts /
// splits and writes LDS with 16-byte stores, 8 consecutive lanes writing one
// chunk of 8 consecutive tile rows (bank-conflict free under lds_k<S, 1>).
template <int MODE, typename TIN>
struct CrossLoad {
  using Pol = CrossPol<MODE>;
  using S = typename Pol::S;
  using V = typename std::conditional<MODE == CR_F64, double, float>::type;
  static constexpr int RK = Pol::RK, LDK = RK + Pol::PAD, HR = RK / 2;
  TIN v[HR];  // raw loads: conversion happens in commit(), after the MFMAs
  int64_t r0_, rows_;

  __device__ inline void fetch(const TIN* __restrict__ src, const void*, int64_t ld, int64_t r0, int64_t rows,
                               int64_t c0, int64_t cols) {
    const int col = threadIdx.x % CR_TM;
    const int rh = (threadIdx.x / CR_TM) * HR;
    r0_ = r0;
    rows_ = rows;
    if (r0 + RK <= rows && c0 + CR_TM <= cols) {
      // interior tile (workgroup-uniform test): uniform base, 32-bit offsets
      const TIN* base = src + r0 * ld + c0;
      const int ld32 = (int)ld;
#pragma unroll
      for (int e = 0; e < HR; ++e) v[e] = base[(rh + e) * ld32 + col];
      return;
    }
    const int64_t gc = c0 + col;
    const bool cok = gc < cols;
#pragma unroll
    for (int e = 0; e < HR; ++e) {
      const int64_t gr = r0 + rh + e;
      v[e] = (cok && gr < rows) ? src[gr * ld + gc] : (TIN)0;
    }
  }

  // scale: segment sign (applied before rounding to the MFMA operand type);
  // roww: optional per-row weights
  __device__ inline void commit(S* hi, S* lo, V scale, const double* roww) const {
    const int col = threadIdx.x % CR_TM;
    const int rh = (threadIdx.x / CR_TM) * HR;
    V x[HR];
#pragma unroll
    for (int e = 0; e < HR; ++e) x[e] = scale * (V)todbl(v[e]);
    if (roww) {
#pragma unroll
      for (int e = 0; e < HR; ++e) {
        const int64_t gr = r0_ + rh + e;
        if (gr < rows_) x[e] = (V)((double)x[e] * roww[gr]);
      }
    }
    S* dh = hi + col * LDK;
    if constexpr (MODE == CR_F64) {
#pragma unroll
      for (int e = 0; e < HR; e += 2) *reinterpret_cast<f64x2*>(dh + rh + e) = f64x2{x[e], x[e + 1]};
    } else if constexpr (MODE == CR_F32) {
#pragma unroll
      for (int e = 0; e < HR; e += 4)
        *reinterpret_cast<f32x4*>(dh + rh + e) = f32x4{x[e], x[e + 1], x[e + 2], x[e + 3]};
    } else {
#pragma unroll
      for (int e = 0; e < HR; e += 8) {
        bf16x8 h, l;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const __bf16 hq = (__bf16)x[e + q];
          h[q] = hq;
          if constexpr (MODE == CR_BF16X3) l[q] = (__bf16)(x[e + q] - (float)hq);
        }
        const int pos = col * LDK + lds_k<S, 1>(col, rh + e);
        *reinterpret_cast<bf16x8*>(hi + pos) = h;
        if constexpr (MODE == CR_BF16X3) *reinterpret_cast<bf16x8*>(lo + pos) = l;
      }
    }
  }
};

// Loader of pre-split operands: thread t owns the column pair 2*(t%64),
// +1 and rows 8*(t/64) .. +7 of the k-step, so each global load is one
// 4-byte pair per lane and one 256-byte row segment per wave.  commit()
// transposes the pairs into two 8-row column vectors per plane (byte
// permutes only) and writes them in an order that keeps every 8 lanes on 4
// even and 4 odd tile rows (lanes with (t/4) odd write their odd column
// first), which is bank-conflict free under lds_k<S, 1>.  A -1 sign flips
// sign bits, exact in both planes.
template <>
struct CrossLoad<CR_BF16X3, SplitBF16> {
  using S = __bf16;
  using V = float;
  static constexpr int RK = CrossPol<CR_BF16X3>::RK, LDK = RK + CrossPol<CR_BF16X3>::PAD, RG = 8;
  static_assert(RK == 4 * RG && CR_TM == 128, "four 8-row chunks x 64 column pairs");
  uint32_t h[RG], l[RG];
  uint32_t cmask;  // columns of the pair inside the matrix (per 16-bit half)
  int nrow;        // rows of this thread's 8 inside the segment

  // Interior tiles load from a uniform base with 32-bit offsets.  Elsewhere
  // every load is still unconditional from an in-bounds address (rows past
  // the end re-read the last row, a pair past the last column re-reads
  // column 0; a pair straddling the last column stays inside the padded row,
  // whose length is even) and commit() zeroes what lies outside, so no
  // branch sits between the loads.
  __device__ inline void fetch(const SplitBF16* src, const void* src_lo, int64_t ld, int64_t r0, int64_t rows,
                               int64_t c0, int64_t cols) {
    const __bf16* H = reinterpret_cast<const __bf16*>(src);
    const __bf16* L = reinterpret_cast<const __bf16*>(src_lo);
    const int cp = threadIdx.x & 63, rg = (threadIdx.x >> 6) * RG;
    if (r0 + RK <= rows && c0 + CR_TM <= cols) {
      const int64_t b = r0 * ld + c0;
      const int ld32 = (int)ld;
      cmask = 0xffffffffu;
      nrow = RG;
#pragma unroll
      for (int e = 0; e < RG; ++e) {
        const int off = (rg + e) * ld32 + 2 * cp;
        h[e] = *reinterpret_cast<const uint32_t*>(H + b + off);
        l[e] = *reinterpret_cast<const uint32_t*>(L + b + off);
      }
      return;
    }
    const int64_t gc = c0 + 2 * cp;
    cmask = gc + 1 < cols ? 0xffffffffu : (gc < cols ? 0x0000ffffu : 0u);
    const int64_t cc = gc < cols ? gc : 0;
    const int64_t rleft = rows - (r0 + rg);
    nrow = rleft >= RG ? RG : (rleft > 0 ? (int)rleft : 0);
#pragma unroll
    for (int e = 0; e < RG; ++e) {
      int64_t gr = r0 + rg + e;
      gr = gr < rows ? gr : rows - 1;
      const int64_t off = gr * ld + cc;
      h[e] = *reinterpret_cast<const uint32_t*>(H + off);
      l[e] = *reinterpret_cast<const uint32_t*>(L + off);
    }
  }

  __device__ inline void commit(S* hi, S* lo, V scale, const double*) const {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const int cp = threadIdx.x & 63, rg = (threadIdx.x >> 6) * RG;
    const uint32_t flip = scale < 0 ? 0x80008000u : 0u;
    const bool oddf = (cp >> 2) & 1;  // write the odd column first
    // byte selectors of v_perm over {b : a}: low halves (column 2p) are
    // bytes 0,1,4,5; high halves (column 2p+1) bytes 2,3,6,7
    const uint32_t sa = oddf ? 0x07060302u : 0x05040100u, sb = oddf ? 0x05040100u : 0x07060302u;
    u32x4 ha, hb, la, lb;
#pragma unroll
    for (int k = 0; k < RG / 2; ++k) {
      const uint32_t ma = 2 * k < nrow ? cmask : 0u, mb = 2 * k + 1 < nrow ? cmask : 0u;
      const uint32_t a = (h[2 * k] ^ flip) & ma, b = (h[2 * k + 1] ^ flip) & mb;
      ha[k] = __builtin_amdgcn_perm(b, a, sa);
      hb[k] = __builtin_amdgcn_perm(b, a, sb);
      const uint32_t c = (l[2 * k] ^ flip) & ma, d = (l[2 * k + 1] ^ flip) & mb;
      la[k] = __builtin_amdgcn_perm(d, c, sa);
      lb[k] = __builtin_amdgcn_perm(d, c, sb);
    }
    const int ca = 2 * cp + (oddf ? 1 : 0), cb = 2 * cp + (oddf ? 0 : 1);
    const int pa = ca * LDK + lds_k<S, 1>(ca, rg), pb = cb * LDK + lds_k<S, 1>(cb, rg);
    *reinterpret_cast<u32x4*>(hi + pa) = ha;
    *reinterpret_cast<u32x4*>(lo + pa) = la;
    *reinterpret_cast<u32x4*>(hi + pb) = hb;
    *reinterpret_cast<u32x4*>(lo + pb) = lb;
  }
};

// Row range of split `split` within a segment of `rows` rows.
template <int RK>
__device__ inline void split_range(int64_t rows, int splits, int split, int64_t& rbeg, int64_t& rend) {
  int64_t chunk = (rows + splits - 1) / splits;
  chunk = (chunk + RK - 1) / RK * RK;
  rbeg = (int64_t)split * chunk;
  rend = rbeg + chunk < rows ? rbeg + chunk : rows;
}

template <int MODE, typename TX, typename TY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void cross_kernel(CrossArgs a) {
  using Pol = CrossPol<MODE>;
  using S = typename Pol::S;
  using frag_t = typename Pol::frag_t;
  using acc_t = typename Pol::acc_t;
  using outer_t = typename Pol::outer_t;
  constexpr int RK = Pol::RK, LDK = RK + Pol::PAD, KI = Pol::KI, KE = Pol::KE;
  constexpr int NL = (MODE == CR_BF16X3) ? 2 : 1;  // hi (+ lo) tiles
  constexpr int TILE = CR_TM * LDK;
  constexpr int STAGE = 2 * NL * TILE;

  extern __shared__ __align__(16) unsigned char smem[];
  S* const buf0 = reinterpret_cast<S*>(smem);
  S* const buf1 = buf0 + STAGE;

  // XCD-aware map: workgroup L runs on XCD L % 8.  The work items
  // u = split * ntiles + tile are dealt to the XCDs in contiguous blocks of
  // N/8, so the tiles one XCD runs at a time share their row splits, which
  // are read from HBM once and then served from that XCD's L2.
  int split, tile;
  {
    const int L = blockIdx.x;
    const int N = a.ntiles * a.splits;
    const int u = (N % 8 == 0) ? (L & 7) * (N >> 3) + (L >> 3) : L;
    split = u / a.ntiles;
    tile = u % a.ntiles;
  }
  int ti, tj;
  cross_tile_ij(a, tile, ti, tj);
  const int64_t i0 = (int64_t)ti * CR_TM, j0 = (int64_t)tj * CR_TN;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int fr = lane & 15, kg = (lane >> 4) * KE;
  // fragment rows are 16m + fr, so their chunk swizzle depends on fr only
  // (for bf16 KI == RK: kk is 0 and the swizzle is exact; fp32/fp64 k is unswizzled)
  const int koff = lds_k<S, 1>(fr, kg);

  // Two-level accumulation (MFMA registers flushed into `outer` every FLUSH
  // rows) only where the fp32 rounding chain matters: the fp32 and split
  // bf16 Gram modes.  Single bf16 products (weight gradients) and fp64 use
  // the MFMA accumulators alone.
  constexpr bool TWO = Pol::FLUSH < (1 << 20);
  acc_t inner[4][4];
  outer_t outer[TWO ? 4 : 1][TWO ? 4 : 1][TWO ? 4 : 1];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) inner[i][j] = acc_t{0, 0, 0, 0};
  if constexpr (TWO) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) outer[i][j][r] = 0;
  }

  CrossLoad<MODE, TX> lx;
  CrossLoad<MODE, TY> ly;
  // flattened (segment, k-step) iteration with one-stage prefetch
  int s = 0;
  int64_t rbeg = 0, rend = 0, r0 = 0;
  auto next_seg = [&](int from) {
    for (s = from; s < a.nseg; ++s) {
      split_range<RK>(a.seg[s].rows, a.splits, split, rbeg, rend);
      if (rbeg < rend) { r0 = rbeg; return; }
    }
  };
  next_seg(0);
  if (s >= a.nseg) goto write_out;
  lx.fetch(reinterpret_cast<const TX*>(a.seg[s].x), a.seg[s].xlo, a.seg[s].ldx, r0, rend, i0, a.m);
  ly.fetch(reinterpret_cast<const TY*>(a.seg[s].y), a.seg[s].ylo, a.seg[s].ldy, r0, rend, j0, a.n);
  lx.commit(buf0, buf0 + TILE, (typename CrossLoad<MODE, TX>::V)a.seg[s].sign, a.roww);
  ly.commit(buf0 + NL * TILE, buf0 + NL * TILE + TILE, 1, nullptr);
  __syncthreads();
  {
    int stage = 0, since = 0;
    (void)since;
    while (true) {
      // position of the next k-step
      const bool seg_last = r0 + RK >= rend;
      int ns = s;
      int64_t nr0 = r0 + RK;
      if (seg_last) {
        next_seg(s + 1);
        ns = s;
        nr0 = r0;
      }
      const bool more = ns < a.nseg;
      if (more) {
        lx.fetch(reinterpret_cast<const TX*>(a.seg[ns].x), a.seg[ns].xlo, a.seg[ns].ldx, nr0, rend, i0, a.m);
        ly.fetch(reinterpret_cast<const TY*>(a.seg[ns].y), a.seg[ns].ylo, a.seg[ns].ldy, nr0, rend, j0, a.n);
      }
      const S* xs = stage ? buf1 : buf0;
      const S* ys = xs + NL * TILE;
#pragma unroll
      for (int kk = 0; kk < RK; kk += KI) {
        frag_t yh[4], yl[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int off = (wj + j * 16 + fr) * LDK + koff + kk;
          yh[j] = *reinterpret_cast<const frag_t*>(ys + off);
          if constexpr (MODE == CR_BF16X3) yl[j] = *reinterpret_cast<const frag_t*>(ys + TILE + off);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int off = (wi + i * 16 + fr) * LDK + koff + kk;
          const frag_t xh = *reinterpret_cast<const frag_t*>(xs + off);
          if constexpr (MODE == CR_BF16X3) {
            const frag_t xl = *reinterpret_cast<const frag_t*>(xs + TILE + off);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              inner[i][j] = Pol::mfma(xl, yh[j], inner[i][j]);
              inner[i][j] = Pol::mfma(xh, yl[j], inner[i][j]);
              inner[i][j] = Pol::mfma(xh, yh[j], inner[i][j]);
            }
          } else {
#pragma unroll
            for (int j = 0; j < 4; ++j) inner[i][j] = Pol::mfma(xh, yh[j], inner[i][j]);
          }
        }
      }
      if constexpr (TWO) {
        since += RK;
        if (since >= Pol::FLUSH || !more) {
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
              for (int r = 0; r < 4; ++r) outer[i][j][r] += (outer_t)inner[i][j][r];
              inner[i][j] = acc_t{0, 0, 0, 0};
            }
          since = 0;
        }
      }
      if (!more) break;
      stage ^= 1;
      S* nb = stage ? buf1 : buf0;
      lx.commit(nb, nb + TILE, (typename CrossLoad<MODE, TX>::V)a.seg[ns].sign, a.roww);
      ly.commit(nb + NL * TILE, nb + NL * TILE + TILE, 1, nullptr);
      __syncthreads();
      r0 = nr0;
    }
  }

write_out:
  double* out = a.partial + ((int64_t)split * a.ntiles + tile) * (CR_TM * CR_TN);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int li = wi + i * 16 + Pol::drow(lane, r);
        const int lj = wj + j * 16 + fr;
        if constexpr (TWO) out[li * CR_TN + lj] = (double)outer[i][j][r];
        else out[li * CR_TN + lj] = (double)inner[i][j][r];
      }
}

// G = beta*G + scale * sum_split partial (FP64 sums of FP64 or fp32 partial
// tiles), over 32x32 output blocks.  For the
// symmetric mode only blocks with bi <= bj run; each writes (i,j) and the
// mirrored (j,i) through an LDS transpose, and diagonal blocks take every
// value from the upper triangle, so G is exactly symmetric.  All global
// reads and writes are row-contiguous.
template <typename TP>
__global__ __launch_bounds__(256) void cross_reduce_kernel(const TP* __restrict__ partial, int splits,
                                                            int ntiles, int tiles_n, int symmetric, int bdiag,
                                                            int64_t m, int64_t n, double scale, double beta,
                                                            double* __restrict__ G, int64_t ldg) {
  __shared__ double tr[32][33];
  const int64_t nbj = (n + 31) / 32;
  int64_t bi, bj;
  if (symmetric) {
    int64_t t = blockIdx.x, rowlen = nbj;
    bi = 0;
    while (t >= rowlen) { t -= rowlen; ++bi; --rowlen; }
    bj = bi + t;
  } else {
    bi = blockIdx.x / nbj;
    bj = blockIdx.x % nbj;
  }
  // block-diagonal tiles only: 32 x 32 blocks outside them are not computed
  // (a block lies inside one tile: 32 divides the tile sides)
  if (bdiag > 0 && (bi * 32 / CR_TM) / bdiag != (bj * 32 / CR_TN) / bdiag) return;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t stride = (int64_t)ntiles * CR_TM * CR_TN;
  // the thread's 4 rows share one loop over the splits, 8 splits at a time:
  // 32 independent loads in flight per thread, each row summed in split order
  // (rows past the output read the tile's first element and are not stored:
  // conditional loads would compile to a branch and a wait per load)
  const TP* p[4];
  bool ok[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = bi * 32 + ty + 8 * q, j = bj * 32 + tx;
    ok[q] = i < m && j < n;
    const int64_t ti = ok[q] ? i / CR_TM : 0, tj = ok[q] ? j / CR_TN : 0;
    const int64_t tile = symmetric ? ti * tiles_n - ti * (ti - 1) / 2 + (tj - ti)
                         : bdiag > 0 ? (ti / bdiag) * bdiag * bdiag + (ti % bdiag) * bdiag + (tj % bdiag)
                                     : ti * tiles_n + tj;
    p[q] = partial + tile * (CR_TM * CR_TN) + (ok[q] ? (i % CR_TM) * CR_TN + (j % CR_TN) : 0);
  }
  double val[4] = {0.0, 0.0, 0.0, 0.0};
  int sp = 0;
  for (; sp + 8 <= splits; sp += 8) {
    TP t[4][8];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < 8; ++u) t[q][u] = p[q][(sp + u) * stride];  // valid even for rows past the output
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int u = 0; u < 8; ++u) val[q] += (double)t[q][u];
  }
  for (; sp < splits; ++sp) {
#pragma unroll
    for (int q = 0; q < 4; ++q) val[q] += (double)p[q][sp * stride];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) tr[ty + 8 * q][tx] = val[q];
  if (!symmetric) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = bi * 32 + ty + 8 * q, j = bj * 32 + tx;
      if (i < m && j < n) {
        double* g = G + i * ldg + j;
        *g = (beta == 0.0 ? 0.0 : beta * *g) + scale * val[q];
      }
    }
    return;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int row = ty + 8 * q;
    const int64_t i = bi * 32 + row, j = bj * 32 + tx;
    if (i < m && j < n) {
      const double v = (bi == bj && row > tx) ? tr[tx][row] : val[q];
      double* g = G + i * ldg + j;
      *g = (beta == 0.0 ? 0.0 : beta * *g) + scale * v;
    }
    if (bi != bj) {
      // mirrored block: G[bj*32 + row][bi*32 + tx] = value at (bi*32+tx, bj*32+row)
      const int64_t mi = bj * 32 + row, mj = bi * 32 + tx;
      if (mi < n && mj < m) {
        double* g = G + mi * ldg + mj;
        *g = (beta == 0.0 ? 0.0 : beta * *g) + scale * tr[tx][row];
      }
    }
  }
}

template <int MODE>
constexpr size_t cross_lds_bytes() {  // two stages of X and Y tiles (hi [+ lo])
  using Pol = CrossPol<MODE>;
  constexpr int NL = (MODE == CR_BF16X3) ? 2 : 1;
  return 2 * sizeof(typename Pol::S) * (size_t)(Pol::RK + Pol::PAD) * CR_TM * NL * 2;
}
