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

constexpr int RG_MAXCH = 12;
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
};

template <typename T> __device__ inline T fneg(T v) { return -v; }
template <> __device__ inline bf16x8 fneg<bf16x8>(bf16x8 v) {
  typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
  u16x8 u = __builtin_bit_cast(u16x8, v);
  u ^= (unsigned short)0x8000;
  return __builtin_bit_cast(bf16x8, u);
}

// Cooperative load of an R x RK tile (rows r0.., k0..) of a row-major source
// into LDS [R][LDK], converting TIN -> S.  Out-of-range rows / k are zero.
template <int NTHR, int R, int RK, int LDK, typename TIN, typename S>
__device__ inline void load_rows(S* __restrict__ lds, const TIN* __restrict__ src, int64_t ld,
                                 int64_t r0, int64_t nrows, int64_t k0, int64_t kmax) {
  constexpr int VE = 16 / sizeof(TIN);
  static_assert(RK % VE == 0, "tile k must be a multiple of the vector width");
  constexpr int VPR = RK / VE;
  constexpr int NV = R * VPR;
  for (int v = threadIdx.x; v < NV; v += NTHR) {
    const int row = v / VPR;
    const int col = (v % VPR) * VE;
    const int64_t gr = r0 + row;
    const int64_t gk = k0 + col;
    TIN vals[VE];
    if (gr < nrows && gk + VE <= kmax) {
      typedef TIN vec_t __attribute__((ext_vector_type(VE)));
      vec_t t = *reinterpret_cast<const vec_t*>(src + gr * ld + gk);
#pragma unroll
      for (int e = 0; e < VE; ++e) vals[e] = t[e];
    } else {
#pragma unroll
      for (int e = 0; e < VE; ++e)
        vals[e] = (gr < nrows && gk + e < kmax) ? src[gr * ld + gk + e] : (TIN)0;
    }
    S* dst = lds + row * LDK + col;
#pragma unroll
    for (int e = 0; e < VE; ++e) dst[e] = (S)vals[e];
  }
}

template <int P, typename TIN, class Plan>
__global__ __launch_bounds__(256) void rowgemm_kernel(RowArgs a) {
  using PR = Prec<P>;
  using S = typename PR::store_t;
  using M = typename PR::math_t;
  using frag_t = typename PR::frag_t;
  using acc_t = typename PR::acc_t;
  constexpr int RK = PR::RK, LDK = RK + PR::PAD, KI = PR::KI, KE = PR::KE;
  constexpr int CI = Plan::CI, NW = Plan::NW, NA = Plan::NA, NT = Plan::NT;
  constexpr int FB = Plan::FB, PB = Plan::PB, WPN = Plan::WPN, WPP = Plan::WPP;
  static_assert(WPN * WPP == 4, "four waves per workgroup");
  constexpr int TN = WPN * FB * 16, TP = WPP * PB * 16;

  extern __shared__ __align__(16) unsigned char smem[];
  S* ws = reinterpret_cast<S*>(smem);    // [NW][TN][LDK]
  S* xs = ws + NW * TN * LDK;            // [CI][TP][LDK]

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wn = wave % WPN, wp = wave / WPN;
  const int64_t p0 = (int64_t)blockIdx.x * TP;
  const int64_t f0 = (int64_t)blockIdx.y * TN;

  acc_t acc[NA][FB][PB];
#pragma unroll
  for (int s = 0; s < NA; ++s)
#pragma unroll
    for (int i = 0; i < FB; ++i)
#pragma unroll
      for (int j = 0; j < PB; ++j) acc[s][i][j] = acc_t{0, 0, 0, 0};

  const int fr = lane & 15, kg = (lane >> 4) * KE;
  for (int64_t k0 = 0; k0 < a.k; k0 += RK) {
#pragma unroll
    for (int s = 0; s < NW; ++s)
      load_rows<256, TN, RK, LDK, S, S>(ws + s * TN * LDK, reinterpret_cast<const S*>(a.w[s]), a.ldw[s],
                                        f0, a.nout, k0, a.k);
#pragma unroll
    for (int c = 0; c < CI; ++c)
      load_rows<256, TP, RK, LDK, TIN, S>(xs + c * TP * LDK, reinterpret_cast<const TIN*>(a.x[c]), a.ldx[c],
                                          p0, a.npts, k0, a.k);
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < RK; kk += KI) {
      frag_t wf[NW][FB], xf[CI][PB];
#pragma unroll
      for (int s = 0; s < NW; ++s)
#pragma unroll
        for (int i = 0; i < FB; ++i)
          wf[s][i] = *reinterpret_cast<const frag_t*>(ws + (s * TN + wn * FB * 16 + i * 16 + fr) * LDK + kk + kg);
#pragma unroll
      for (int c = 0; c < CI; ++c)
#pragma unroll
        for (int j = 0; j < PB; ++j)
          xf[c][j] = *reinterpret_cast<const frag_t*>(xs + (c * TP + wp * PB * 16 + j * 16 + fr) * LDK + kk + kg);
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
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FB; ++i)
#pragma unroll
    for (int j = 0; j < PB; ++j) {
      const int64_t p = p0 + wp * PB * 16 + j * 16 + fr;
      const int64_t fb = f0 + wn * FB * 16 + i * 16;
      if (p >= a.npts) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t f = fb + PR::drow(lane, r);
        if (f >= a.nout) continue;
        M v[NA];
#pragma unroll
        for (int s = 0; s < NA; ++s) v[s] = (M)acc[s][i][j][r];
        Plan::template epilogue<P>(a, p, f, v);
      }
    }
}

// ---------------------------------------------------------------------------
// Plans
// ---------------------------------------------------------------------------
template <int D> struct TileFor;
template <> struct TileFor<2> { static constexpr int FB = 4, PB = 1, WPN = 2, WPP = 2; };
template <> struct TileFor<3> { static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2; };

// Hidden layer forward: Z = omega*(X W^T + b) (all channels), Y = f-jet(Z).
template <int D, int ACT> struct PlanHiddenFwd {
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  static constexpr int ta[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tc[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tw[10] = {0};
  static constexpr bool tneg[10] = {false};
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    using S = typename Prec<P>::store_t;
    M u[C], y[C];
    const M om = (M)a.omega;
    if (a.bias) v[0] += reinterpret_cast<const M*>(a.bias)[f];
#pragma unroll
    for (int c = 0; c < C; ++c) u[c] = om * v[c];
    act_jet_forward<D, ACT, M>(u, y);
    S* z = reinterpret_cast<S*>(a.out0);
    S* o = reinterpret_cast<S*>(a.out1);
    const int64_t off = p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < C; ++c) {
      z[c * a.cstride + off] = cvt<S>(u[c]);
      o[c * a.cstride + off] = cvt<S>(y[c]);
    }
  }
};

// Output layer, raw jets: O = X W^T (+ b on the value channel), field_t out.
template <int D> struct PlanOutJets {
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  static constexpr int ta[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tc[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tw[10] = {0};
  static constexpr bool tneg[10] = {false};
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    if (a.bias) v[0] += reinterpret_cast<const M*>(a.bias)[f];
    M* o = reinterpret_cast<M*>(a.out0);
    const int64_t off = p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < C; ++c) o[c * a.cstride + off] = v[c];
  }
};

// 2D output layer with the velocity / vorticity lift (ELC:35-38):
//   u = psi_y, v = -psi_x  (mass fields),  omega = -(psi_xx + psi_yy).
// Inputs: x[0]=gx, x[1]=gy, x[2]=hxx, x[3]=hyy of the last hidden jet.
struct PlanOutFields2 {
  static constexpr int CI = 4, NW = 1, NA = 3, NT = 4;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {0, 1, 2, 2};
  static constexpr int tc[NT] = {1, 0, 2, 3};
  static constexpr int tw[NT] = {0, 0, 0, 0};
  static constexpr bool tneg[NT] = {false, true, true, true};
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    M* mass = reinterpret_cast<M*>(a.out0);
    M* stiff = reinterpret_cast<M*>(a.out1);
    if (a.layout == 0) {
      mass[(2 * p) * a.ldo + f] = v[0];
      mass[(2 * p + 1) * a.ldo + f] = v[1];
    } else {
      mass[p * a.ldo + f] = v[0];
      mass[(a.npts + p) * a.ldo + f] = v[1];
    }
    stiff[p * a.ldo2 + f] = v[2];
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
struct PlanOutFields3 {
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
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    M* U = reinterpret_cast<M*>(a.out0);
    M* Dg = reinterpret_cast<M*>(a.out1);
    if (a.layout == 0) {
#pragma unroll
      for (int c = 0; c < 3; ++c) U[(3 * p + c) * a.ldo + f] = v[c];
#pragma unroll
      for (int c = 0; c < 9; ++c) Dg[(9 * p + c) * a.ldo2 + f] = v[3 + c];
    } else {
#pragma unroll
      for (int c = 0; c < 3; ++c) U[(c * a.npts + p) * a.ldo + f] = v[c];
#pragma unroll
      for (int c = 0; c < 9; ++c) Dg[(c * a.npts + p) * a.ldo2 + f] = v[3 + c];
    }
  }
};

// Hidden layer reverse: A = A'_{l} W_l (all channels) then the transposed
// activation chain of layer l-1 with its tape Z:  out = omega * f^T(Z, A).
// The host passes W_l^T so the kernel still contracts X . W^T.
template <int D, int ACT> struct PlanHiddenBwd {
  static constexpr int C = Jet2Layout<D>::C;
  static constexpr int CI = C, NW = 1, NA = C, NT = C;
  static constexpr int FB = TileFor<D>::FB, PB = TileFor<D>::PB, WPN = TileFor<D>::WPN, WPP = TileFor<D>::WPP;
  static constexpr int ta[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tc[10] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};
  static constexpr int tw[10] = {0};
  static constexpr bool tneg[10] = {false};
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    using S = typename Prec<P>::store_t;
    const S* z = reinterpret_cast<const S*>(a.tape);
    M u[C], b[C];
    const int64_t toff = p * a.ldt + f;
#pragma unroll
    for (int c = 0; c < C; ++c) u[c] = tomath<M>(z[c * a.tstride + toff]);
    act_jet_reverse<D, ACT, M>(u, v, b);
    S* o = reinterpret_cast<S*>(a.out0);
    const M om = (M)a.omega;
    const int64_t off = p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < C; ++c) o[c * a.cstride + off] = cvt<S>(om * b[c]);
  }
};

// 2D output-layer reverse fused with the lift transpose (ELC:35-38 transposed)
// and the last activation transpose.  Inputs: x[0]=u rows, x[1]=v rows of
// the mass fields, x[2]=omega rows of the stiffness fields.  Weights:
// w[0] = Bm = 2w W^T dL/dM, w[1] = Bs = 2w W^T dL/dS  ([width][K]).
//   adj gx = -(v . Bm), adj gy = u . Bm, adj hxx = adj hyy = -(omega . Bs)
struct PlanOutBwd2 {
  static constexpr int CI = 3, NW = 2, NA = 3, NT = 3;
  static constexpr int FB = 4, PB = 2, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {1, 0, 2};
  static constexpr int tc[NT] = {0, 1, 2};
  static constexpr int tw[NT] = {0, 0, 1};
  static constexpr bool tneg[NT] = {false, true, true};
};

template <int ACT> struct PlanOutBwd2A : PlanOutBwd2 {
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    using S = typename Prec<P>::store_t;
    M adj[6] = {(M)0, v[0], v[1], v[2], (M)0, v[2]};
    const S* z = reinterpret_cast<const S*>(a.tape);
    M u[6], b[6];
    const int64_t toff = p * a.ldt + f;
#pragma unroll
    for (int c = 0; c < 6; ++c) u[c] = tomath<M>(z[c * a.tstride + toff]);
    act_jet_reverse<2, ACT, M>(u, adj, b);
    S* o = reinterpret_cast<S*>(a.out0);
    const M om = (M)a.omega;
    const int64_t off = p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < 6; ++c) o[c * a.cstride + off] = cvt<S>(om * b[c]);
  }
};

// 3D output-layer reverse: the transpose of PlanOutFields3 term by term.
// Inputs x[0..2] = U component rows, x[3..11] = grad-u rows (3a+j).
// Weights w[0..2] = 2w W_c^T dL/dM, w[3..5] = 2w W_c^T dL/dS.
// Accumulators = adjoints of the last hidden jet channels g0..g2, h0..h5.
template <int ACT> struct PlanOutBwd3 {
  static constexpr int CI = 12, NW = 6, NA = 9, NT = 24;
  static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2;
  static constexpr int ta[NT] = {PlanOutFields3::tc[0], PlanOutFields3::tc[1], PlanOutFields3::tc[2],
                                 PlanOutFields3::tc[3], PlanOutFields3::tc[4], PlanOutFields3::tc[5],
                                 PlanOutFields3::tc[6], PlanOutFields3::tc[7], PlanOutFields3::tc[8],
                                 PlanOutFields3::tc[9], PlanOutFields3::tc[10], PlanOutFields3::tc[11],
                                 PlanOutFields3::tc[12], PlanOutFields3::tc[13], PlanOutFields3::tc[14],
                                 PlanOutFields3::tc[15], PlanOutFields3::tc[16], PlanOutFields3::tc[17],
                                 PlanOutFields3::tc[18], PlanOutFields3::tc[19], PlanOutFields3::tc[20],
                                 PlanOutFields3::tc[21], PlanOutFields3::tc[22], PlanOutFields3::tc[23]};
  static constexpr int tc[NT] = {PlanOutFields3::ta[0], PlanOutFields3::ta[1], PlanOutFields3::ta[2],
                                 PlanOutFields3::ta[3], PlanOutFields3::ta[4], PlanOutFields3::ta[5],
                                 PlanOutFields3::ta[6], PlanOutFields3::ta[7], PlanOutFields3::ta[8],
                                 PlanOutFields3::ta[9], PlanOutFields3::ta[10], PlanOutFields3::ta[11],
                                 PlanOutFields3::ta[12], PlanOutFields3::ta[13], PlanOutFields3::ta[14],
                                 PlanOutFields3::ta[15], PlanOutFields3::ta[16], PlanOutFields3::ta[17],
                                 PlanOutFields3::ta[18], PlanOutFields3::ta[19], PlanOutFields3::ta[20],
                                 PlanOutFields3::ta[21], PlanOutFields3::ta[22], PlanOutFields3::ta[23]};
  // mass fields (terms 0..5) use w[c], stiffness fields (terms 6..23) w[3+c]
  static constexpr int tw[NT] = {PlanOutFields3::tw[0], PlanOutFields3::tw[1], PlanOutFields3::tw[2],
                                 PlanOutFields3::tw[3], PlanOutFields3::tw[4], PlanOutFields3::tw[5],
                                 3 + PlanOutFields3::tw[6], 3 + PlanOutFields3::tw[7], 3 + PlanOutFields3::tw[8],
                                 3 + PlanOutFields3::tw[9], 3 + PlanOutFields3::tw[10], 3 + PlanOutFields3::tw[11],
                                 3 + PlanOutFields3::tw[12], 3 + PlanOutFields3::tw[13], 3 + PlanOutFields3::tw[14],
                                 3 + PlanOutFields3::tw[15], 3 + PlanOutFields3::tw[16], 3 + PlanOutFields3::tw[17],
                                 3 + PlanOutFields3::tw[18], 3 + PlanOutFields3::tw[19], 3 + PlanOutFields3::tw[20],
                                 3 + PlanOutFields3::tw[21], 3 + PlanOutFields3::tw[22], 3 + PlanOutFields3::tw[23]};
  static constexpr bool tneg[NT] = {false, true, false, true, false, true, false, true, false, true, false, true,
                                    false, true, false, true, false, true, false, true, false, true, false, true};
  template <int P> __device__ static inline void epilogue(const RowArgs& a, int64_t p, int64_t f,
                                                         typename Prec<P>::math_t* v) {
    using M = typename Prec<P>::math_t;
    using S = typename Prec<P>::store_t;
    M adj[10];
    adj[0] = (M)0;
#pragma unroll
    for (int c = 0; c < 9; ++c) adj[1 + c] = v[c];
    const S* z = reinterpret_cast<const S*>(a.tape);
    M u[10], b[10];
    const int64_t toff = p * a.ldt + f;
#pragma unroll
    for (int c = 0; c < 10; ++c) u[c] = tomath<M>(z[c * a.tstride + toff]);
    act_jet_reverse<3, ACT, M>(u, adj, b);
    S* o = reinterpret_cast<S*>(a.out0);
    const M om = (M)a.omega;
    const int64_t off = p * a.ldo + f;
#pragma unroll
    for (int c = 0; c < 10; ++c) o[c * a.cstride + off] = cvt<S>(om * b[c]);
  }
};

template <int P, class Plan>
constexpr size_t rowgemm_lds_bytes() {
  using PR = Prec<P>;
  constexpr int LDK = PR::RK + PR::PAD;
  constexpr int TN = Plan::WPN * Plan::FB * 16, TP = Plan::WPP * Plan::PB * 16;
  return sizeof(typename PR::store_t) * (size_t)LDK * (Plan::NW * TN + Plan::CI * TP);
}
