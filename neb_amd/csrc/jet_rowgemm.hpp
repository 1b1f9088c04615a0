// Multi-channel row GEMM for packed second-order spatial jets.
//
// For a block of points and a block of output features it accumulates NA
// "accumulator sets", each a signed sum of terms  X_c . W_s^T  where X_c is
// one jet channel of the input ([points][k], k contiguous) and W_s one weight
// slice ([features][k]).  Every set shares the MFMA result layout, so lane l
// holds the same (point, feature) pairs in every set: the activation jet
// chain, the curl / Laplacian lifts and their transposes are then plain
// per-lane arithmetic in the epilogue, with no cross-lane traffic and no HBM
// round trip for the unactivated GEMM output.
//
// MFMA orientation: A = weight fragment (rows = features), B = input
// fragment (cols = points), so lane l holds point l%16 and four features
// (rows drow(l, r), r = 0..3) of each 16x16 block.
//
// Reference operations replaced (paths relative to the reference root):
//   packed Jet2 linear layer   HB2P:1005-1036, VJ2:20-27   (forward plans)
//   fused sine / unary Jet2    FUSED:3-8, HB2P:1038-1040   (forward epilogue)
//   linear transpose           HB2P:1115-1144, VJ2:28-30   (backward plans)
//   unary Jet2 transpose       HB2P:1147-1174, FUSED:9-19  (backward epilogue)
//   lifts / their transposes   ELC:35-40, VJ2:93-106, NCC:97-109, VCORE:32-38
// where HB2P = experiments/2d/_backend/periodic_core/vortex_merger/src/
// eigenfluid/gen/vortex_merger2d/hardcoded_backend.py, VJ2/NCC/VCORE/FUSED =
// experiments/metaball/_backend/flow/cases/metaball/{velocity_trace_jet2,
// native_cupy_core,velocity/core,velocity/fused}.py, ELC = experiments/2d/
// _backend/periodic_core/vortex_merger/canonical_gauge/scripts/
// evaluate_large_comparison.py.
#pragma once
#include "neb_common.hpp"
#include <type_traits>

constexpr int RG_MAXCH = 20;
constexpr int RG_MAXW = 6;

struct RowArgs {
  const void* x[RG_MAXCH];  // input channel bases ([points][k], row stride ldx[c])
  int64_t ldx[RG_MAXCH];
  const void* w[RG_MAXW];   // weight slices ([features][k], row stride ldw[s])
  int64_t ldw[RG_MAXW];
  int64_t npts, nout, k;
  const void* bias;         // math_t[nout] or null
  double omega;
  void* out0;               // plan-specific outputs
  void* out1;
  int64_t ldo, cstride;     // output row stride and packed channel stride
  void* out2;
  int64_t ldo2;
  const void* tape;         // pre-activation jet tape (packed, store_t)
  int64_t ldt, tstride;
  int layout;               // fields: 0 = point-major rows, 1 = component-major
  void* out0lo;             // split fields: bf16 lo planes beside out0 / out1
  void* out1lo;
  const void* env;          // envelope jets, math_t [C][npts] (PlanOutJetsT<D, true>)
  int64_t kwin;             // > 0: block-diagonal weights with kwin x kwin blocks; a
                            // feature tile contracts only the k of its blocks
};

template <typename T> __device__ inline T fneg(T v) { return -v; }
template <> __device__ inline bf16x8 fneg<bf16x8>(bf16x8 v) {
  typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
  u16x8 u = __builtin_bit_cast(u16x8, v);
  u ^= (unsigned short)0x8000;
  return __builtin_bit_cast(bf16x8, u);
}

// Joint loader of NSRC tiles (channels or weight slices), each R rows x RK
// k-values of a row-major source, into LDS [NSRC][R][LDK].  fetch() issues
// the 16-byte global loads into registers (so they overlap the MFMAs of the
// previous stage); commit() converts TIN -> S and writes LDS with one vector
// store per load.  Rows >= nrows and k >= kmax read as zero.
template <int NSRC, int R, int RK, int LDK, typename TIN, typename S>
struct TileLoad {
  static constexpr int VE = 16 / sizeof(TIN);
  static constexpr int VPR = RK / VE;
  static constexpr int PER = R * VPR;
  static constexpr int NV = NSRC * PER;
  static constexpr int NVT = (NV + 255) / 256;
  static_assert(RK % VE == 0, "tile k must be a multiple of the vector width");
  typedef TIN vin __attribute__((ext_vector_type(VE)));
  typedef S vout __attribute__((ext_vector_type(VE)));
  typedef int i32x4 __attribute__((ext_vector_type(4)));
  vin v[NVT];

  __device__ inline void fetch(const void* const* src, const int64_t* ld, int64_t r0, int64_t nrows, int64_t k0,
                               int64_t kmax) {
#pragma unroll
    for (int t = 0; t < NVT; ++t) {
      const int idx = threadIdx.x + 256 * t;
      vin val = __builtin_bit_cast(vin, i32x4{0, 0, 0, 0});
      if (NV % 256 == 0 || idx < NV) {
        const int s = idx / PER, rem = idx % PER;
        const int row = rem / VPR, col = (rem % VPR) * VE;
        const int64_t gr = r0 + row, gk = k0 + col;
        // select the source with compile-time indices only (a dynamic index
        // into the kernel-argument arrays would force a private copy)
        const void* sp = src[0];
        int64_t sld = ld[0];
#pragma unroll
        for (int q = 1; q < NSRC; ++q)
          if (s == q) { sp = src[q]; sld = ld[q]; }
        if (gr < nrows) {
          const TIN* p = reinterpret_cast<const TIN*>(sp) + gr * sld + gk;
          if (gk + VE <= kmax) {
            val = *reinterpret_cast<const vin*>(p);
          } else {
#pragma unroll
            for (int e = 0; e < VE; ++e)
              if (gk + e < kmax) val[e] = p[e];
          }
        }
      }
      v[t] = val;
    }
  }

  __device__ inline void commit(S* lds) const {
#pragma unroll
    for (int t = 0; t < NVT; ++t) {
      const int idx = threadIdx.x + 256 * t;
      if (NV % 256 == 0 || idx < NV) {
        const int s = idx / PER, rem = idx % PER;
        const int row = rem / VPR, col = (rem % VPR) * VE;
        vout o;
#pragma unroll
        for (int e = 0; e < VE; ++e) o[e] = (S)v[t][e];
        *reinterpret_cast<vout*>(lds + (s * R + row) * LDK + lds_k<S, 2>(row, col)) = o;
      }
    }
  }
};

// Group stores / loads of GE consecutive features (16-byte vectors).
// NT: non-temporal stores (streamed past the caches: outputs no kernel of
// the launch reads back, so they do not evict inputs other tiles reuse)
template <typename T, int GE, typename M, bool NT = false>
__device__ inline void st_group(T* dst, const M* v, int nvalid) {
  constexpr int PV = (int)(16 / sizeof(T)) < GE ? (int)(16 / sizeof(T)) : GE;
  typedef T pv __attribute__((ext_vector_type(PV)));
  if (nvalid >= GE) {
#pragma unroll
    for (int c = 0; c < GE; c += PV) {
      pv t;
#pragma unroll
      for (int e = 0; e < PV; ++e) t[e] = cvt<T>(v[c + e]);
      if constexpr (NT)
        __builtin_nontemporal_store(t, reinterpret_cast<pv*>(dst + c));
      else
        *reinterpret_cast<pv*>(dst + c) = t;
    }
  } else {
#pragma unroll
    for (int e = 0; e < GE; ++e)
      if (e < nvalid) dst[e] = cvt<T>(v[e]);
  }
}

// Field row store at element offset `off`: the math type as is, or (SPLIT)
// hi = bf16(x) into `out` and lo = bf16(x - hi) into `out_lo`.  x - hi is
// exact in fp32, so hi + lo carries 16 significant bits of x.
template <bool SPLIT, int GE, typename M, bool NT = false>
__device__ inline void st_field(void* out, void* out_lo, int64_t off, const M* v, int nv) {
  if constexpr (SPLIT) {
    float h[GE], l[GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      h[e] = (float)(__bf16)(float)v[e];
      l[e] = (float)v[e] - h[e];
    }
    st_group<__bf16, GE, float, NT>(reinterpret_cast<__bf16*>(out) + off, h, nv);
    st_group<__bf16, GE, float, NT>(reinterpret_cast<__bf16*>(out_lo) + off, l, nv);
  } else {
    st_group<M, GE, M, NT>(reinterpret_cast<M*>(out) + off, v, nv);
  }
}

template <typename T, int GE, typename M>
__device__ inline void ld_group(const T* src, M* v, int nvalid) {
  constexpr int PV = (int)(16 / sizeof(T)) < GE ? (int)(16 / sizeof(T)) : GE;
  typedef T pv __attribute__((ext_vector_type(PV)));
  if (nvalid >= GE) {
#pragma unroll
    for (int c = 0; c < GE; c += PV) {
      const pv t = *reinterpret_cast<const pv*>(src + c);
#pragma unroll
      for (int e = 0; e < PV; ++e) v[c + e] = tomath<M>(t[e]);
    }
  } else {
#pragma unroll
    for (int e = 0; e < GE; ++e) v[e] = e < nvalid ? tomath<M>(src[e]) : (M)0;
  }
}

template <int P, class Plan>
constexpr size_t rowgemm_stage_elems() {
  using PR = Prec<P>;
  constexpr int LDK = PR::RK + PR::PAD;
  constexpr int TN = Plan::WPN * Plan::FB * 16, TP = Plan::WPP * Plan::PB * 16;
  return (size_t)LDK * (Plan::NW * TN + Plan::CI * TP);
}

template <int P, class Plan>
constexpr bool rowgemm_double_buffered() {
  return 2 * rowgemm_stage_elems<P, Plan>() * sizeof(typename Prec<P>::store_t) <= 96 * 1024;
}

template <int P, class Plan>
constexpr size_t rowgemm_lds_bytes() {
  return (rowgemm_double_buffered<P, Plan>() ? 2 : 1) * rowgemm_stage_elems<P, Plan>() *
         sizeof(typename Prec<P>::store_t);
}

// Waves per SIMD the register-staged kernel is compiled for: two (256
// registers per lane) unless a plan's accumulators, operand fragments and
// staged loads need more -- those get one wave per SIMD (512 registers)
// instead of spilling to scratch (RowWaves specialisations after the plans).
template <int P, class Plan> struct RowWaves { static constexpr int value = P == PREC_F64 ? 1 : 2; };

template <int P, typename TIN, class Plan>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RowWaves<P, Plan>::value))) void rowgemm_kernel(RowArgs a) {
  using PR = Prec<P>;
  using S = typename PR::store_t;
  using M = typename PR::math_t;
  using frag_t = typename PR::frag_t;
  using acc_t = typename PR::acc_t;
  constexpr int RK = PR::RK, LDK = RK + PR::PAD, KI = PR::KI, KE = PR::KE;
  constexpr int CI = Plan::CI, NW = Plan::NW, NA = Plan::NA, NT = Plan::NT;
  constexpr int FB = Plan::FB, PB = Plan::PB, WPN = Plan::WPN, WPP = Plan::WPP;
  static_assert(WPN * WPP == 4, "four waves per workgroup");
  static_assert(FB % 2 == 0, "8-feature epilogue groups pair feature blocks");
  constexpr int TN = WPN * FB * 16, TP = WPP * PB * 16;
  constexpr int WT = TN * LDK, XT = TP * LDK;
  constexpr int STAGE = NW * WT + CI * XT;
  constexpr bool DB = rowgemm_double_buffered<P, Plan>();

  extern __shared__ __align__(16) unsigned char smem[];
  S* buf0 = reinterpret_cast<S*>(smem);
  S* buf1 = DB ? buf0 + STAGE : buf0;

  // XCD-aware tile map: workgroup L runs on XCD L % 8; all feature tiles of
  // one point block go to the same XCD so its input rows are read from HBM
  // once and served from that XCD's L2 afterwards.
  const int64_t npb = (a.npts + TP - 1) / TP, nfb = (a.nout + TN - 1) / TN;
  int64_t pb, fbk;
  {
    const int64_t L = blockIdx.x;
    if (npb % 8 == 0) {
      const int64_t q = L >> 3;
      pb = (L & 7) + 8 * (q / nfb);
      fbk = q % nfb;
    } else {
      pb = L / nfb;
      fbk = L % nfb;
    }
  }
  const int64_t p0 = pb * TP, f0 = fbk * TN;
  // contraction window [kb, ke): all of k, or with block-diagonal weights
  // (a.kwin) the k of the blocks that features f0 .. f0+TN-1 belong to
  int64_t kb = 0, ke = a.k;
  if (a.kwin > 0) {
    kb = f0 / a.kwin * a.kwin;
    const int64_t e = ((f0 + TN - 1) / a.kwin + 1) * a.kwin;
    ke = e < a.k ? e : a.k;
  }

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wn = wave % WPN, wp = wave / WPN;

  acc_t acc[NA][FB][PB];
#pragma unroll
  for (int s = 0; s < NA; ++s)
#pragma unroll
    for (int i = 0; i < FB; ++i)
#pragma unroll
      for (int j = 0; j < PB; ++j) acc[s][i][j] = acc_t{0, 0, 0, 0};

  // Two register stages of prefetch: while stage kt computes from LDS, the
  // loads of stages kt+1 and kt+2 are both in flight.
  TileLoad<NW, TN, RK, LDK, S, S> lw[2];
  TileLoad<CI, TP, RK, LDK, TIN, S> lx[2];
  const int fr = lane & 15, kg = (lane >> 4) * KE;
  // Feature permutation: result element (block i, row drow(lane, r)) of lane
  // group g = lane/16 is feature phi(g, i, r) of the wave's 16*FB features,
  //   GE = 8 (2-byte outputs): phi = 32*(i/2) + 8*g + 4*(i%2) + r
  //   GE = 4 (wider outputs):  phi = 16*i + 4*g + r
  // so that in every epilogue store instruction the four lane groups of a
  // point write adjacent 16-byte chunks (whole 32-byte sectors).
  constexpr int GE = Plan::template out_bytes<P>() == 2 ? 8 : 4;
  const int rg = PR::rowg(fr), rr = PR::rowr(fr);
  const int nk = (int)((ke - kb + RK - 1) / RK);

  // the second in-flight stage only where its registers fit beside the
  // accumulators (raw operands no wider than the LDS type)
  constexpr int NVEC = TileLoad<NW, TN, RK, LDK, S, S>::NVT + TileLoad<CI, TP, RK, LDK, TIN, S>::NVT;
  constexpr bool DEEP = sizeof(S) == 2 && sizeof(TIN) == 2 && NA * FB * PB <= 24 && NVEC <= 6;
  lw[0].fetch(a.w, a.ldw, f0, a.nout, kb, ke);
  lx[0].fetch(a.x, a.ldx, p0, a.npts, kb, ke);
  if (DEEP && nk > 1) {
    lw[1].fetch(a.w, a.ldw, f0, a.nout, kb + RK, ke);
    lx[1].fetch(a.x, a.ldx, p0, a.npts, kb + RK, ke);
  }
  lw[0].commit(buf0);
  lx[0].commit(buf0 + NW * WT);
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur_set = DEEP ? (kt & 1) : 1;
    S* cur = (DB && (kt & 1)) ? buf1 : buf0;
    if (!DEEP && kt + 1 < nk) {  // single in-flight stage: fetch kt+1 into set 0
      const int64_t k1 = kb + (int64_t)(kt + 1) * RK;
      lw[0].fetch(a.w, a.ldw, f0, a.nout, k1, ke);
      lx[0].fetch(a.x, a.ldx, p0, a.npts, k1, ke);
    }
    if (DEEP && kt + 2 < nk) {
      const int64_t k2 = kb + (int64_t)(kt + 2) * RK;
      if (cur_set == 0) {
        lw[0].fetch(a.w, a.ldw, f0, a.nout, k2, ke);
        lx[0].fetch(a.x, a.ldx, p0, a.npts, k2, ke);
      } else {
        lw[1].fetch(a.w, a.ldw, f0, a.nout, k2, ke);
        lx[1].fetch(a.x, a.ldx, p0, a.npts, k2, ke);
      }
    }
    const S* ws = cur;
    const S* xs = cur + NW * WT;
#pragma unroll
    for (int kk = 0; kk < RK; kk += KI) {
      frag_t wf[NW][FB], xf[CI][PB];
#pragma unroll
      for (int i = 0; i < FB; ++i) {
        const int row = wn * FB * 16 + (GE == 8 ? 32 * (i >> 1) + 8 * rg + 4 * (i & 1) : 16 * i + 4 * rg) + rr;
        const int off = row * LDK + lds_k<S, 2>(row, kk + kg);
#pragma unroll
        for (int s = 0; s < NW; ++s) wf[s][i] = *reinterpret_cast<const frag_t*>(ws + s * WT + off);
      }
#pragma unroll
      for (int j = 0; j < PB; ++j) {
        const int row = wp * PB * 16 + j * 16 + fr;
        const int off = row * LDK + lds_k<S, 2>(row, kk + kg);
#pragma unroll
        for (int c = 0; c < CI; ++c) xf[c][j] = *reinterpret_cast<const frag_t*>(xs + c * XT + off);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int i = 0; i < FB; ++i) {
          const frag_t wv = Plan::tneg[t] ? fneg(wf[Plan::tw[t]][i]) : wf[Plan::tw[t]][i];
#pragma unroll
          for (int j = 0; j < PB; ++j)
            acc[Plan::ta[t]][i][j] = PR::mfma(wv, xf[Plan::tc[t]][j], acc[Plan::ta[t]][i][j]);
        }
    }
    if (kt + 1 < nk) {
      if (!DB) __syncthreads();
      S* nxt = (DB && !(kt & 1)) ? buf1 : buf0;
      if (DEEP && cur_set == 0) {  // stage kt+1 lives in register set (kt+1)&1 (set 0 when not DEEP)
        lw[1].commit(nxt);
        lx[1].commit(nxt + NW * WT);
      } else {
        lw[0].commit(nxt);
        lx[0].commit(nxt + NW * WT);
      }
    }
    __syncthreads();
  }

  const int g = lane >> 4;
#pragma unroll
  for (int j = 0; j < PB; ++j) {
    const int64_t p = p0 + wp * PB * 16 + j * 16 + fr;
#pragma unroll
    for (int q = 0; q < FB * 4 / GE; ++q) {
      const int64_t f = f0 + wn * FB * 16 + (GE == 8 ? 32 * q + 8 * g : 16 * q + 4 * g);
      const int64_t rem = a.nout - f;
      if (p < a.npts && rem > 0) {
        const int nvalid = rem >= GE ? GE : (int)rem;
        M v[NA][GE];
#pragma unroll
        for (int s = 0; s < NA; ++s)
#pragma unroll
          for (int e = 0; e < GE; ++e)
            v[s][e] = GE == 8 ? (M)acc[s][2 * q + (e >> 2)][j][e & 3] : (M)acc[s][q][j][e];
        Plan::template epilogue<P, GE>(a, p, f, nvalid, v);
      }
    }
  }
}

// The same plans when every weight row holds one nonzero, in column
// f / kdiag (the output layer of a block of per-mode networks: output
// column f = m kdiag + j belongs to mode m, whose weights read only mode m's
// field cotangents), the weights given as those nonzeros, w_s[f]:
// acc[s](p, f) = sum over the plan's terms t with ta[t] = s of
// (-)w_tw[t][f] x_tc[t][p][f / kdiag], then the plan's epilogue.  One thread
// per (point, group of GE features); the contraction is a few loads, the
// epilogue (tape reads, jet chain, stores) is the bandwidth.  kdiag: a
// multiple of GE.
template <int P, typename TIN, class Plan>
__global__ __launch_bounds__(256) void rowdiag_kernel(RowArgs a, int64_t kdiag) {
  using PR = Prec<P>;
  using S = typename PR::store_t;
  using M = typename PR::math_t;
  constexpr int GE = Plan::template out_bytes<P>() == 2 ? 8 : 4;
  const int64_t ngr = (a.nout + GE - 1) / GE;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= a.npts * ngr) return;
  const int64_t p = idx / ngr, f = idx % ngr * GE;
  const int64_t k = f / kdiag;
  const int nvalid = a.nout - f >= GE ? GE : (int)(a.nout - f);
  M xin[Plan::CI];
#pragma unroll
  for (int c = 0; c < Plan::CI; ++c) xin[c] = tomath<M>(reinterpret_cast<const TIN*>(a.x[c])[p * a.ldx[c] + k]);
  M v[Plan::NA][GE];
#pragma unroll
  for (int s = 0; s < Plan::NA; ++s)
#pragma unroll
    for (int e = 0; e < GE; ++e) v[s][e] = (M)0;
  M wv[Plan::NW][GE];
#pragma unroll
  for (int s = 0; s < Plan::NW; ++s) ld_group<S, GE>(reinterpret_cast<const S*>(a.w[s]) + f, wv[s], nvalid);
#pragma unroll
  for (int t = 0; t < Plan::NT; ++t)
#pragma unroll
    for (int e = 0; e < GE; ++e)
      v[Plan::ta[t]][e] += (Plan::tneg[t] ? -wv[Plan::tw[t]][e] : wv[Plan::tw[t]][e]) * xin[Plan::tc[t]];
  Plan::template epilogue<P, GE>(a, p, f, nvalid, v);
}

// ---------------------------------------------------------------------------
// Plans.  epilogue<P, GE>(a, p, f, nvalid, v): v[s][e] is accumulator set s
// at point p, feature f+e (e < nvalid <= GE).
// ---------------------------------------------------------------------------
template <int D> struct TileFor;
template <> struct TileFor<2> { static constexpr int FB = 4, PB = 1, WPN = 2, WPP = 2; };
template <> struct TileFor<3> { static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2; };
template <> struct TileFor<JET1_2D> { static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2; };

#define IDENTITY_TERMS                                             \
  static constexpr int ta[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};    \
  static constexpr int tc[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};    \
  static constexpr int tw[10] = {0};                               \
  static constexpr bool tneg[10] = {false};

// Hidden layer forward: Z = omega*(X W^T + b) (all channels), Y = f-jet(Z).
template <int D, int ACT> struct PlanHiddenFwd {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  IDENTITY_TERMS
  // The value channel takes the bias (a.bias, fp32 [nout] or null); a
  // caller that has added it to v[0] already passes BIAS_ADDED.
  static constexpr bool BIAS_CH0 = true;
  template <int P, int GE, typename M, bool BIAS_ADDED = false>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    using S = typename Prec<P>::store_t;
    using L = Jet2Layout<D>;
    const M om = (M)a.omega;
    const M* bias = reinterpret_cast<const M*>(a.bias);
    M u[C][GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      const M b = (!BIAS_ADDED && bias && e < nv) ? bias[f + e] : (M)0;
      u[0][e] = BIAS_ADDED ? om * v[0][e] : om * (v[0][e] + b);
#pragma unroll
      for (int c = 1; c < C; ++c) u[c][e] = om * v[c][e];
    }
    const int64_t off = p * a.ldo + f;
    S* z = reinterpret_cast<S*>(a.out0) + off;
    S* y = reinterpret_cast<S*>(a.out1) + off;
#pragma unroll
    for (int c = 0; c < C; ++c) st_group<S, GE>(z + c * a.cstride, u[c], nv);
    M f1[GE], f2[GE], t[GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) Act<ACT, M>::eval(u[0][e], t[e], f1[e], f2[e]);
    st_group<S, GE>(y, t, nv);
#pragma unroll
    for (int c = 0; c < L::S; ++c) {
#pragma unroll
      for (int e = 0; e < GE; ++e) t[e] = f1[e] * u[1 + c][e];
      st_group<S, GE>(y + (1 + c) * a.cstride, t, nv);
    }
#pragma unroll
    for (int q = 0; q < L::P; ++q) {
#pragma unroll
      for (int e = 0; e < GE; ++e)
        t[e] = f1[e] * u[1 + L::S + q][e] + f2[e] * u[1 + L::pi(q)][e] * u[1 + L::pj(q)][e];
      st_group<S, GE>(y + (1 + L::S + q) * a.cstride, t, nv);
    }
  }
};

// Split-precision operands: the input jets and the weights are each held
// as two bf16 planes, x = x_hi + x_lo and W = W_hi + W_lo, and every product
// is W_lo x_hi + W_hi x_lo + W_hi x_hi on bf16 MFMA with fp32 accumulation
// (W_lo x_lo is below fp32 rounding), so the contraction carries ~16
// significant bits of each operand instead of 8.  Inputs x[0..CH-1] are the
// hi planes and x[CH..2CH-1] the lo planes of the same channels; w[0] = W_hi,
// w[1] = W_lo.  Each base term (accumulator a, channel c, sign) becomes the
// three terms (a, c, W_lo), (a, c + CH, W_hi), (a, c, W_hi): small products
// first.

// Term tables of the split hidden layers: channel c into accumulator c.
template <int D> struct HiddenSplitTerms;
template <> struct HiddenSplitTerms<2> {
  static constexpr int ta[18] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5};
  static constexpr int tc[18] = {0, 6, 0, 1, 7, 1, 2, 8, 2, 3, 9, 3, 4, 10, 4, 5, 11, 5};
  static constexpr int tw[18] = {1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
  static constexpr bool tneg[18] = {false};
};
template <> struct HiddenSplitTerms<3> {
  static constexpr int ta[30] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8,
                                 9, 9, 9};
  static constexpr int tc[30] = {0, 10, 0, 1, 11, 1, 2, 12, 2, 3, 13, 3, 4, 14, 4, 5, 15, 5, 6, 16, 6, 7, 17, 7,
                                 8, 18, 8, 9, 19, 9};
  static constexpr int tw[30] = {1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0,
                                 1, 0, 0};
  static constexpr bool tneg[30] = {false};
};
template <> struct HiddenSplitTerms<JET1_2D> {
  static constexpr int ta[9] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
  static constexpr int tc[9] = {0, 3, 0, 1, 4, 1, 2, 5, 2};
  static constexpr int tw[9] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
  static constexpr bool tneg[9] = {false};
};

// Hidden layer forward on split operands: Z = omega*(X W^T + b) and
// Y = f-jet(Z) as PlanHiddenFwd.  Y is stored as hi / lo bf16 planes (out1 /
// out1lo: the next layer's split input); Z as its hi plane (out0; the bf16
// tape of the reverse pass), plus its lo plane when out0lo is set.
// 3D split hidden tile: 4 waves across 128 features x 16 points (each wave
// 32 x 16), a 36 KB stage that double-buffers at two workgroups per CU
template <int D> struct SplitHiddenTile : TileFor<D> {};
template <> struct SplitHiddenTile<3> { static constexpr int FB = 2, PB = 1, WPN = 4, WPP = 1; };

template <int D, int ACT> struct PlanHiddenFwdSplit : HiddenSplitTerms<D> {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 2; }
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = 2 * C, NW = 2, NA = C, NT = 3 * C;
  using TL = SplitHiddenTile<D>;
  static constexpr int FB = TL::FB, PB = TL::PB, WPN = TL::WPN, WPP = TL::WPP;
  // the value channel takes the bias (a.bias, fp32 [nout] or null); a caller
  // that has added it to v[0] already passes BIAS_ADDED
  static constexpr bool BIAS_CH0 = true;
  template <int P, int GE, typename M, bool BIAS_ADDED = false>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    using L = Jet2Layout<D>;
    const M om = (M)a.omega;
    const M* bias = reinterpret_cast<const M*>(a.bias);
    M u[C][GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      const M b = (!BIAS_ADDED && bias && e < nv) ? bias[f + e] : (M)0;
      u[0][e] = om * (v[0][e] + b);
#pragma unroll
      for (int c = 1; c < C; ++c) u[c][e] = om * v[c][e];
    }
    const int64_t off = p * a.ldo + f;
    if (a.out0lo) {
#pragma unroll
      for (int c = 0; c < C; ++c) st_field<true, GE, M>(a.out0, a.out0lo, off + c * a.cstride, u[c], nv);
    } else {
#pragma unroll
      for (int c = 0; c < C; ++c) st_group<__bf16, GE>(reinterpret_cast<__bf16*>(a.out0) + off + c * a.cstride, u[c], nv);
    }
    M f1[GE], f2[GE], t[GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) Act<ACT, M>::eval(u[0][e], t[e], f1[e], f2[e]);
    st_field<true, GE, M>(a.out1, a.out1lo, off, t, nv);
#pragma unroll
    for (int c = 0; c < L::S; ++c) {
#pragma unroll
      for (int e = 0; e < GE; ++e) t[e] = f1[e] * u[1 + c][e];
      st_field<true, GE, M>(a.out1, a.out1lo, off + (1 + c) * a.cstride, t, nv);
    }
#pragma unroll
    for (int q = 0; q < L::P; ++q) {
#pragma unroll
      for (int e = 0; e < GE; ++e)
        t[e] = f1[e] * u[1 + L::S + q][e] + f2[e] * u[1 + L::pi(q)][e] * u[1 + L::pj(q)][e];
      st_field<true, GE, M>(a.out1, a.out1lo, off + (1 + L::S + q) * a.cstride, t, nv);
    }
  }
};

// Output layer, raw jets: O = X W^T (+ b on the value channel), field_t out;
// with ENV the stored jets are the product e O with the per-point envelope
// jet e (jet_product, VJ2:47-50; NCC:94-95 field = e * raw), one envelope
// per point shared by every output feature.
template <int D, bool ENV = false> struct PlanOutJetsT {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::math_t); }
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  IDENTITY_TERMS
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    const M* bias = reinterpret_cast<const M*>(a.bias);
    if (bias)
#pragma unroll
      for (int e = 0; e < GE; ++e) v[0][e] += e < nv ? bias[f + e] : (M)0;
    if constexpr (ENV) {
      using MT = typename Prec<P>::math_t;
      const MT* env = reinterpret_cast<const MT*>(a.env) + p;
      M ej[C];
#pragma unroll
      for (int c = 0; c < C; ++c) ej[c] = (M)env[c * a.npts];
#pragma unroll
      for (int e = 0; e < GE; ++e) {
        M x[C], y[C];
#pragma unroll
        for (int c = 0; c < C; ++c) x[c] = v[c][e];
        jet_product<D, M>(ej, x, y);
#pragma unroll
        for (int c = 0; c < C; ++c) v[c][e] = y[c];
      }
    }
    M* o = reinterpret_cast<M*>(a.out0) + p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < C; ++c) st_group<M, GE>(o + c * a.cstride, v[c], nv);
  }
};
template <int D> using PlanOutJets = PlanOutJetsT<D, false>;

// 2D output layer with the velocity / vorticity lift (ELC:35-38):
//   u = psi_y, v = -psi_x  (mass fields),  omega = -(psi_xx + psi_yy).
// Inputs: x[0]=gx, x[1]=gy, x[2]=hxx, x[3]=hyy of the last hidden jet.
template <bool SPLIT, bool STREAM = false> struct PlanOutFields2T {
  template <int P> __host__ __device__ static constexpr int out_bytes() {
    return SPLIT ? 2 : sizeof(typename Prec<P>::math_t);
  }
  static constexpr int CI = 4, NW = 1, NA = 3, NT = 4;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 1, 2, 2};
  static constexpr int tc[NT] = {1, 0, 2, 3};
  static constexpr int tw[NT] = {0, 0, 0, 0};
  static constexpr bool tneg[NT] = {false, true, true, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    const int64_t r0 = a.layout == 0 ? 2 * p : p;
    const int64_t r1 = a.layout == 0 ? 2 * p + 1 : a.npts + p;
    st_field<SPLIT, GE, M, STREAM>(a.out0, a.out0lo, r0 * a.ldo + f, v[0], nv);
    st_field<SPLIT, GE, M, STREAM>(a.out0, a.out0lo, r1 * a.ldo + f, v[1], nv);
    st_field<SPLIT, GE, M, STREAM>(a.out1, a.out1lo, p * a.ldo2 + f, v[2], nv);
  }
};
using PlanOutFields2 = PlanOutFields2T<false>;

// 2D output layer + lift on split operands (as above): inputs x[0..3]
// = hi planes of gx, gy, hxx, hyy, x[4..7] their lo planes; w[0] = W_hi,
// w[1] = W_lo; the fields are stored as hi / lo bf16 planes.
struct PlanOutFields2Split {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 2; }
  using B = PlanOutFields2T<true, true>;
  static constexpr int CI = 8, NW = 2, NA = 3, NT = 12;
  static constexpr int FB = B::FB, PB = 1, WPN = B::WPN, WPP = B::WPP;
  // base terms of PlanOutFields2T (CH = 4): u += gy, v -= gx, omega -= hxx, omega -= hyy
  static constexpr int ta[NT] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 2, 2, 2};
  static constexpr int tc[NT] = {1, 5, 1, 0, 4, 0, 2, 6, 2, 3, 7, 3};
  static constexpr int tw[NT] = {1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false, true, true, true, true, true, true, true, true, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    B::template epilogue<P, GE, M>(a, p, f, nv, v);
  }
};

// 2D output layer with the velocity lift and the Hessian stiffness of the
// causal velocity peel (HB2P:2935-3060): mass fields u = psi_y, v = -psi_x;
// stiffness rows (psi_xx, sqrt2 psi_xy, psi_yy), whose Gram is the Hessian
// Frobenius form with the pair metric (1, 2, 1) (HB2P:3022).
// Inputs: x[0]=gx, x[1]=gy, x[2]=hxx, x[3]=hxy, x[4]=hyy.
constexpr float SQRT2F = 1.41421356237309504880f;
template <bool SPLIT, bool STREAM = false> struct PlanOutFields2HT {
  template <int P> __host__ __device__ static constexpr int out_bytes() {
    return SPLIT ? 2 : sizeof(typename Prec<P>::math_t);
  }
  static constexpr int CI = 5, NW = 1, NA = 5, NT = 5;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 1, 2, 3, 4};
  static constexpr int tc[NT] = {1, 0, 2, 3, 4};
  static constexpr int tw[NT] = {0, 0, 0, 0, 0};
  static constexpr bool tneg[NT] = {false, true, false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    const int64_t r0 = a.layout == 0 ? 2 * p : p;
    const int64_t r1 = a.layout == 0 ? 2 * p + 1 : a.npts + p;
    st_field<SPLIT, GE, M, STREAM>(a.out0, a.out0lo, r0 * a.ldo + f, v[0], nv);
    st_field<SPLIT, GE, M, STREAM>(a.out0, a.out0lo, r1 * a.ldo + f, v[1], nv);
#pragma unroll
    for (int e = 0; e < GE; ++e) v[3][e] *= (M)(P == PREC_F64 ? 1.41421356237309504880 : SQRT2F);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int64_t row = a.layout == 0 ? 3 * p + r : r * a.npts + p;
      st_field<SPLIT, GE, M, STREAM>(a.out1, a.out1lo, row * a.ldo2 + f, v[2 + r], nv);
    }
  }
};

// The same on split operands: x[0..4] hi planes of gx, gy, hxx, hxy, hyy,
// x[5..9] their lo planes; w[0] = W_hi, w[1] = W_lo; three products per base
// term (W_lo x_hi, W_hi x_lo, W_hi x_hi), hi / lo field planes out.
struct PlanOutFields2HSplit {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 2; }
  using B = PlanOutFields2HT<true, true>;
  static constexpr int CI = 10, NW = 2, NA = 5, NT = 15;
  static constexpr int FB = B::FB, PB = 1, WPN = B::WPN, WPP = B::WPP;
  static constexpr int ta[NT] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 4, 4, 4};
  static constexpr int tc[NT] = {1, 6, 1, 0, 5, 0, 2, 7, 2, 3, 8, 3, 4, 9, 4};
  static constexpr int tw[NT] = {1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0, 1, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false, true, true, true, false, false, false,
                                    false, false, false, false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    B::template epilogue<P, GE, M>(a, p, f, nv, v);
  }
};

// 2D output layer with the streamfunction lift of the Dirichlet Rayleigh
// peel (HB2P:3383-3474): mass rows psi (one per point), stiffness rows
// (psi_x, psi_y) (two per point, like the velocity rows).  No output bias:
// the peel removes each mode's mean over the points (the periodic gauge,
// HB2P:3475-3479), which cancels it.
// Inputs: x[0]=v, x[1]=gx, x[2]=gy of the last hidden jet.
template <bool SPLIT, bool STREAM = false> struct PlanOutFields2ST {
  template <int P> __host__ __device__ static constexpr int out_bytes() {
    return SPLIT ? 2 : sizeof(typename Prec<P>::math_t);
  }
  static constexpr int CI = 3, NW = 1, NA = 3, NT = 3;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 1, 2};
  static constexpr int tc[NT] = {0, 1, 2};
  static constexpr int tw[NT] = {0, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    const int64_t r0 = a.layout == 0 ? 2 * p : p;
    const int64_t r1 = a.layout == 0 ? 2 * p + 1 : a.npts + p;
    st_field<SPLIT, GE, M, STREAM>(a.out0, a.out0lo, p * a.ldo + f, v[0], nv);
    st_field<SPLIT, GE, M, STREAM>(a.out1, a.out1lo, r0 * a.ldo2 + f, v[1], nv);
    st_field<SPLIT, GE, M, STREAM>(a.out1, a.out1lo, r1 * a.ldo2 + f, v[2], nv);
  }
};

// The same on split operands: x[0..2] hi planes of v, gx, gy, x[3..5] their
// lo planes; w[0] = W_hi, w[1] = W_lo; hi / lo field planes out.
struct PlanOutFields2SSplit {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 2; }
  using B = PlanOutFields2ST<true, true>;
  static constexpr int CI = 6, NW = 2, NA = 3, NT = 9;
  static constexpr int FB = B::FB, PB = 1, WPN = B::WPN, WPP = B::WPP;
  static constexpr int ta[NT] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
  static constexpr int tc[NT] = {0, 3, 0, 1, 4, 1, 2, 5, 2};
  static constexpr int tw[NT] = {1, 0, 0, 1, 0, 0, 1, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false, false, false, false, false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    B::template epilogue<P, GE, M>(a, p, f, nv, v);
  }
};

// 3D output layer with the curl lift (NCC:97-99, VJ2:93-101).  Vector
// potential A has output rows 3k+c, passed as three weight slices W_c.
// Inputs x[0..2] = grad channels g0..g2, x[3..8] = Hessian pairs h0..h5.
//   u_a      = d_{a+1} A_{a+2} - d_{a+2} A_{a+1}
//   d_j u_a  = d_j d_{a+1} A_{a+2} - d_j d_{a+2} A_{a+1}
struct PairIdx3 {
  __host__ __device__ static constexpr int hid(int a, int b) {
    return a > b ? hid(b, a) : (a == 0 ? b : (a == 1 ? 2 + b : 5));
  }
};
template <bool SPLIT> struct PlanOutFields3T {
  template <int P> __host__ __device__ static constexpr int out_bytes() {
    return SPLIT ? 2 : sizeof(typename Prec<P>::math_t);
  }
  static constexpr int CI = 9, NW = 3, NA = 12, NT = 24;
  static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2;
#define A1(a) ((a + 1) % 3)
#define A2(a) ((a + 2) % 3)
  static constexpr int ta[NT] = {0, 0, 1, 1, 2, 2,
                                 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11};
  static constexpr int tc[NT] = {
      A1(0), A2(0), A1(1), A2(1), A1(2), A2(2),
#define DT(a, j) 3 + PairIdx3::hid(j, A1(a)), 3 + PairIdx3::hid(j, A2(a))
      DT(0, 0), DT(0, 1), DT(0, 2), DT(1, 0), DT(1, 1), DT(1, 2), DT(2, 0), DT(2, 1), DT(2, 2)};
#undef DT
  static constexpr int tw[NT] = {
      A2(0), A1(0), A2(1), A1(1), A2(2), A1(2),
#define DW(a) A2(a), A1(a), A2(a), A1(a), A2(a), A1(a)
      DW(0), DW(1), DW(2)};
#undef DW
#undef A1
#undef A2
  static constexpr bool tneg[NT] = {false, true, false, true, false, true, false, true, false, true, false, true,
                                    false, true, false, true, false, true, false, true, false, true, false, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t r = a.layout == 0 ? 3 * p + c : c * a.npts + p;
      st_field<SPLIT, GE, M>(a.out0, a.out0lo, r * a.ldo + f, v[c], nv);
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
      const int64_t r = a.layout == 0 ? 9 * p + c : c * a.npts + p;
      st_field<SPLIT, GE, M>(a.out1, a.out1lo, r * a.ldo2 + f, v[3 + c], nv);
    }
  }
};
using PlanOutFields3 = PlanOutFields3T<false>;

// 3D output layer + curl lift on split operands: inputs x[0..8] = hi planes
// of g0..g2, h0..h5, x[9..17] their lo planes; w[0..2] = W_hi slices, w[3..5]
// = W_lo slices; each term of PlanOutFields3T expanded as the 2D split plan.
struct PlanOutFields3Split {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 2; }
  using B = PlanOutFields3T<true>;
  static constexpr int CI = 18, NW = 6, NA = 12, NT = 72;
  static constexpr int FB = B::FB, PB = B::PB, WPN = B::WPN, WPP = B::WPP;
  static constexpr int ta[NT] = {0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3,
                                 4, 4, 4, 4, 4, 4, 5, 5, 5, 5, 5, 5, 6, 6, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7,
                                 8, 8, 8, 8, 8, 8, 9, 9, 9, 9, 9, 9, 10, 10, 10, 10, 10, 10, 11, 11, 11, 11, 11, 11};
  static constexpr int tc[NT] = {1, 10, 1, 2, 11, 2, 2, 11, 2, 0, 9, 0, 0, 9, 0, 1, 10, 1, 4, 13, 4, 5, 14, 5,
                                 6, 15, 6, 7, 16, 7, 7, 16, 7, 8, 17, 8, 5, 14, 5, 3, 12, 3, 7, 16, 7, 4, 13, 4,
                                 8, 17, 8, 5, 14, 5, 3, 12, 3, 4, 13, 4, 4, 13, 4, 6, 15, 6, 5, 14, 5, 7, 16, 7};
  static constexpr int tw[NT] = {5, 2, 2, 4, 1, 1, 3, 0, 0, 5, 2, 2, 4, 1, 1, 3, 0, 0, 5, 2, 2, 4, 1, 1,
                                 5, 2, 2, 4, 1, 1, 5, 2, 2, 4, 1, 1, 3, 0, 0, 5, 2, 2, 3, 0, 0, 5, 2, 2,
                                 3, 0, 0, 5, 2, 2, 4, 1, 1, 3, 0, 0, 4, 1, 1, 3, 0, 0, 4, 1, 1, 3, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false, true, true, true, false, false, false, true, true, true,
                                    false, false, false, true, true, true, false, false, false, true, true, true,
                                    false, false, false, true, true, true, false, false, false, true, true, true,
                                    false, false, false, true, true, true, false, false, false, true, true, true,
                                    false, false, false, true, true, true, false, false, false, true, true, true,
                                    false, false, false, true, true, true, false, false, false, true, true, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    B::template epilogue<P, GE, M>(a, p, f, nv, v);
  }
};

// Every split plan's terms are its base plan's, each expanded in the order
// (W_lo x_hi), (W_hi x_lo), (W_hi x_hi): checked at compile time.
template <class Split, int NB, int CH, int NWB, class Base>
constexpr bool split_terms_match() {
  for (int t = 0; t < NB; ++t)
    for (int r = 0; r < 3; ++r) {
      const int q = 3 * t + r;
      if (Split::ta[q] != Base::ta[t] || Split::tneg[q] != Base::tneg[t]) return false;
      if (Split::tc[q] != Base::tc[t] + (r == 1 ? CH : 0)) return false;
      if (Split::tw[q] != Base::tw[t] + (r == 0 ? NWB : 0)) return false;
    }
  return true;
}
struct IdentityTerms10 { IDENTITY_TERMS };
static_assert(split_terms_match<PlanOutFields3Split, 24, 9, 3, PlanOutFields3T<true>>(), "3D split field terms");
static_assert(split_terms_match<PlanOutFields2Split, 4, 4, 1, PlanOutFields2T<true>>(), "2D split field terms");
static_assert(split_terms_match<PlanOutFields2HSplit, 5, 5, 1, PlanOutFields2HT<true>>(), "2D Hessian split terms");
static_assert(split_terms_match<PlanOutFields2SSplit, 3, 3, 1, PlanOutFields2ST<true>>(), "2D stream split terms");
static_assert(split_terms_match<HiddenSplitTerms<2>, 6, 6, 1, IdentityTerms10>(), "2D split hidden terms");
static_assert(split_terms_match<HiddenSplitTerms<3>, 10, 10, 1, IdentityTerms10>(), "3D split hidden terms");

// Shared epilogue of the reverse plans: given the adjoint of the activation
// output (adj[C][GE]), apply the transposed jet chain with the tape Z and
// the omega scale, and store the packed pre-activation adjoint.
template <int P, int D, int ACT, int GE, typename M>
__device__ inline void reverse_epilogue_u(const RowArgs& a, int64_t p, int64_t f, int nv, M (*adj)[GE],
                                          M (*u)[GE]) {
  using S = typename Prec<P>::store_t;
  constexpr int C = Jet2Layout<D>::C;
  M b[C][GE];
  const M om = (M)a.omega;
#pragma unroll
  for (int e = 0; e < GE; ++e) {
    M uc[C], ac[C], bc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) { uc[c] = u[c][e]; ac[c] = adj[c][e]; }
    act_jet_reverse<D, ACT, M>(uc, ac, bc);
#pragma unroll
    for (int c = 0; c < C; ++c) b[c][e] = om * bc[c];
  }
  S* o = reinterpret_cast<S*>(a.out0) + p * a.ldo + f;
#pragma unroll
  for (int c = 0; c < C; ++c) st_group<S, GE>(o + c * a.cstride, b[c], nv);
}

template <int P, int D, int ACT, int GE, typename M>
__device__ inline void reverse_epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*adj)[GE]) {
  using S = typename Prec<P>::store_t;
  constexpr int C = Jet2Layout<D>::C;
  const S* z = reinterpret_cast<const S*>(a.tape) + p * a.ldt + f;
  M u[C][GE];
#pragma unroll
  for (int c = 0; c < C; ++c) ld_group<S, GE>(z + c * a.tstride, u[c], nv);
  reverse_epilogue_u<P, D, ACT, GE, M>(a, p, f, nv, adj, u);
}

// Hidden layer reverse: A = A'_{l} W_l (all channels) then the transposed
// activation chain of layer l-1 with its tape Z:  out = omega * f^T(Z, A).
// The host passes W_l^T so the kernel still contracts X . W^T.
template <int D, int ACT> struct PlanHiddenBwd {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  IDENTITY_TERMS
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    reverse_epilogue<P, D, ACT, GE, M>(a, p, f, nv, v);
  }
  // The same with the tape group already loaded (u = Z at p, f..f+GE).
  // Only tanh holds the tile's tape in registers through the last k-step:
  // the sin chain's epilogue needs more registers than the 48 that costs
  // (with it the 8-wave launch spilled 119 registers in 2D, 203 in 3D)
  static constexpr int TAPE_CH = ACT == ACT_TANH ? C : 0;
  template <int P, int GE, typename M>
  __device__ static inline void epilogue_tape(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE],
                                              M (*u)[GE]) {
    reverse_epilogue_u<P, D, ACT, GE, M>(a, p, f, nv, v, u);
  }
};

// 2D output-layer reverse fused with the lift transpose (ELC:35-38 transposed)
// and the last activation transpose.  Inputs: x[0]=u rows, x[1]=v rows of
// the mass fields, x[2]=omega rows of the stiffness fields.  Weights:
// w[0] = Bm = 2w W^T dL/dM, w[1] = Bs = 2w W^T dL/dS  ([width][K]).
//   adj gx = -(v . Bm), adj gy = u . Bm, adj hxx = adj hyy = -(omega . Bs)
template <int ACT> struct PlanOutBwd2 {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int CI = 3, NW = 2, NA = 3, NT = 3;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {1, 0, 2};
  static constexpr int tc[NT] = {0, 1, 2};
  static constexpr int tw[NT] = {0, 0, 1};
  static constexpr bool tneg[NT] = {false, true, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    M adj[6][GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      adj[0][e] = (M)0; adj[1][e] = v[0][e]; adj[2][e] = v[1][e];
      adj[3][e] = v[2][e]; adj[4][e] = (M)0; adj[5][e] = v[2][e];
    }
    reverse_epilogue<P, 2, ACT, GE, M>(a, p, f, nv, adj);
  }
};

// 2D output-layer reverse for the Hessian stiffness (the transpose of
// PlanOutFields2HT): x[0]=u rows, x[1]=v rows, x[2..4] = the stiffness rows
// (psi_xx, sqrt2 psi_xy, psi_yy); w[0] = Bm, w[1] = Bs.
//   adj gx = -(v . Bm), adj gy = u . Bm,
//   adj hxx = s0 . Bs, adj hxy = sqrt2 (s1 . Bs), adj hyy = s2 . Bs
template <int ACT> struct PlanOutBwd2H {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int CI = 5, NW = 2, NA = 5, NT = 5;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {1, 0, 2, 3, 4};
  static constexpr int tc[NT] = {0, 1, 2, 3, 4};
  static constexpr int tw[NT] = {0, 0, 1, 1, 1};
  static constexpr bool tneg[NT] = {false, true, false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    M adj[6][GE];
    const M r2 = (M)(P == PREC_F64 ? 1.41421356237309504880 : SQRT2F);
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      adj[0][e] = (M)0; adj[1][e] = v[0][e]; adj[2][e] = v[1][e];
      adj[3][e] = v[2][e]; adj[4][e] = r2 * v[3][e]; adj[5][e] = v[4][e];
    }
    reverse_epilogue<P, 2, ACT, GE, M>(a, p, f, nv, adj);
  }
};

// 2D output-layer reverse for the streamfunction lift (the transpose of
// PlanOutFields2ST): x[0] = psi rows, x[1] = psi_x rows, x[2] = psi_y rows;
// w[0] = Bm, w[1] = Bs.
//   adj v = psi . Bm, adj gx = sx . Bs, adj gy = sy . Bs
template <int ACT, int JD = 2> struct PlanOutBwd2S {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int CI = 3, NW = 2, NA = 3, NT = 3;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 1, 2};
  static constexpr int tc[NT] = {0, 1, 2};
  static constexpr int tw[NT] = {0, 1, 1};
  static constexpr bool tneg[NT] = {false, false, false};
  // JD: the tape's jets, second order (2: Hessian adjoints zero) or first
  // order (JET1_2D: value and gradient channels only)
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    constexpr int C = Jet2Layout<JD>::C;
    M adj[C][GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      adj[0][e] = v[0][e]; adj[1][e] = v[1][e]; adj[2][e] = v[2][e];
#pragma unroll
      for (int c = 3; c < C; ++c) adj[c][e] = (M)0;
    }
    reverse_epilogue<P, JD, ACT, GE, M>(a, p, f, nv, adj);
  }
};

// 3D output-layer reverse: the transpose of PlanOutFields3 term by term.
// Inputs x[0..2] = U component rows, x[3..11] = grad-u rows (3a+j).
// Weights w[0..2] = 2w W_c^T dL/dM, w[3..5] = 2w W_c^T dL/dS.
// Accumulators = adjoints of the last hidden jet channels g0..g2, h0..h5.
template <int ACT> struct PlanOutBwd3 {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return sizeof(typename Prec<P>::store_t); }
  static constexpr int CI = 12, NW = 6, NA = 9, NT = 24;
  static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2;
  using F = PlanOutFields3;
  static constexpr int ta[NT] = {F::tc[0], F::tc[1], F::tc[2], F::tc[3], F::tc[4], F::tc[5],
                                 F::tc[6], F::tc[7], F::tc[8], F::tc[9], F::tc[10], F::tc[11],
                                 F::tc[12], F::tc[13], F::tc[14], F::tc[15], F::tc[16], F::tc[17],
                                 F::tc[18], F::tc[19], F::tc[20], F::tc[21], F::tc[22], F::tc[23]};
  static constexpr int tc[NT] = {F::ta[0], F::ta[1], F::ta[2], F::ta[3], F::ta[4], F::ta[5],
                                 F::ta[6], F::ta[7], F::ta[8], F::ta[9], F::ta[10], F::ta[11],
                                 F::ta[12], F::ta[13], F::ta[14], F::ta[15], F::ta[16], F::ta[17],
                                 F::ta[18], F::ta[19], F::ta[20], F::ta[21], F::ta[22], F::ta[23]};
  // mass fields (terms 0..5) use w[c], stiffness fields (terms 6..23) w[3+c]
  static constexpr int tw[NT] = {F::tw[0], F::tw[1], F::tw[2], F::tw[3], F::tw[4], F::tw[5],
                                 3 + F::tw[6], 3 + F::tw[7], 3 + F::tw[8], 3 + F::tw[9], 3 + F::tw[10],
                                 3 + F::tw[11], 3 + F::tw[12], 3 + F::tw[13], 3 + F::tw[14], 3 + F::tw[15],
                                 3 + F::tw[16], 3 + F::tw[17], 3 + F::tw[18], 3 + F::tw[19], 3 + F::tw[20],
                                 3 + F::tw[21], 3 + F::tw[22], 3 + F::tw[23]};
  static constexpr bool tneg[NT] = {false, true, false, true, false, true, false, true, false, true, false, true,
                                    false, true, false, true, false, true, false, true, false, true, false, true};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    M adj[10][GE];
#pragma unroll
    for (int e = 0; e < GE; ++e) {
      adj[0][e] = (M)0;
#pragma unroll
      for (int c = 0; c < 9; ++c) adj[1 + c][e] = v[c][e];
    }
    reverse_epilogue<P, 3, ACT, GE, M>(a, p, f, nv, adj);
  }
};

// Plain GEMM on split operands with an fp32 result (see neb_gemm_split):
// out[p][f] = sum_k x[p][k] w[f][k] with x = x[0] + x[1] and w = w[0] + w[1]
// (hi / lo bf16 planes; the three products of the split hidden layers).
struct PlanGemmSplit {
  template <int P> __host__ __device__ static constexpr int out_bytes() { return 4; }
  static constexpr int CI = 2, NW = 2, NA = 1, NT = 3;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 0, 0};
  static constexpr int tc[NT] = {0, 1, 0};
  static constexpr int tw[NT] = {1, 0, 0};
  static constexpr bool tneg[NT] = {false, false, false};
  template <int P, int GE, typename M>
  __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f, int nv, M (*v)[GE]) {
    st_group<float, GE, M>(reinterpret_cast<float*>(a.out0) + p * a.ldo + f, v[0], nv);
  }
};

// The activation-jet transpose of a layer on its own (see
// neb_jet_act_reverse): the adjoint of the post-activation jet channels
// map.ch[i], given as fp32 columns [i * width, (i+1) * width) of a
// point-major matrix (channels not in the map have a zero adjoint), to the
// adjoint of the scaled pre-activation jet: out = omega * f^T(Z, adj),
// packed [C][n][ldo].  Eight features per thread.
template <int D, int ACT, typename S>
__global__ __launch_bounds__(256) void act_reverse_stacked_kernel(const float* __restrict__ adj, int64_t lda,
                                                                  ChanMap map, const S* __restrict__ z, int64_t ldz,
                                                                  int64_t n, int64_t width, double omega,
                                                                  S* __restrict__ out, int64_t ldo) {
  constexpr int C = Jet2Layout<D>::C;
  const int64_t per = (width + 7) / 8;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n * per) return;
  const int64_t p = idx / per, f0 = (idx % per) * 8;
  const int nv = (int)(width - f0 < 8 ? width - f0 : 8);
  float u[C][8], a[C][8];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    ld_group<S, 8, float>(z + ((int64_t)c * n + p) * ldz + f0, u[c], nv);
#pragma unroll
    for (int e = 0; e < 8; ++e) a[c][e] = 0.0f;
  }
#pragma unroll
  for (int t = 0; t < STACK_MAXCH; ++t) {
    if (t < map.n) {
      const int c = map.ch[t];
      float v[8];
      ld_group<float, 8, float>(adj + p * lda + (int64_t)t * width + f0, v, nv);
#pragma unroll
      for (int cc = 0; cc < C; ++cc)
        if (cc == c) {
#pragma unroll
          for (int e = 0; e < 8; ++e) a[cc][e] += v[e];
        }
    }
  }
  const float om = (float)omega;
  float b[C][8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    float uc[C], ac[C], bc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) { uc[c] = u[c][e]; ac[c] = a[c][e]; }
    act_jet_reverse<D, ACT, float>(uc, ac, bc);
#pragma unroll
    for (int c = 0; c < C; ++c) b[c][e] = om * bc[c];
  }
#pragma unroll
  for (int c = 0; c < C; ++c) st_group<S, 8, float>(out + ((int64_t)c * n + p) * ldo + f0, b[c], nv);
}

// Plans that spilled at two waves per SIMD (measured with
// -Rpass-analysis=kernel-resource-usage): the 3D output layer + curl and its
// transpose, the 2D Hessian-row output layer and its transpose, and the
// sine transposes of the output layer in reduced precision.
template <int P> struct RowWaves<P, PlanOutFields3Split> { static constexpr int value = 1; };
template <int P, bool S> struct RowWaves<P, PlanOutFields3T<S>> { static constexpr int value = 1; };
template <int P, bool S, bool T> struct RowWaves<P, PlanOutFields2HT<S, T>> { static constexpr int value = 1; };
template <int P, int A> struct RowWaves<P, PlanOutBwd2H<A>> { static constexpr int value = 1; };
template <int P, int A> struct RowWaves<P, PlanOutBwd3<A>> { static constexpr int value = P == PREC_BF16 ? 2 : 1; };
template <int P> struct RowWaves<P, PlanOutBwd2<ACT_SIN>> { static constexpr int value = 1; };
