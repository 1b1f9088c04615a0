// extern "C" entry points of libnebhip.so (declared in include/neb_hip.h).
// Argument validation, template dispatch and kernel launches; no allocation.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>

#include "../../include/neb_hip.h"
#include "cross_async.hpp"
#include "crossgemm.hpp"
#include "galerkin.hpp"
#include "jet_rowgemm.hpp"
#include "misc_kernels.hpp"
#include "mode_fields.hpp"
#include "rowgemm_async.hpp"
#include "spd_inverse.hpp"
#include "spectral_terms.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

int hip_status(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(NEB_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  return NEB_OK;
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

size_t elem_size(int dtype) { return dtype == NEB_BF16 ? 2 : (dtype == NEB_F32 ? 4 : 8); }

bool aligned(const void* p, int64_t ld, int dtype) {
  const size_t es = elem_size(dtype);
  return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && ((ld * (int64_t)es) % 16 == 0);
}

template <typename K>
int set_lds(K kernel, size_t bytes) {
  if (bytes > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e != hipSuccess) return fail(NEB_ERR_HIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
  }
  return NEB_OK;
}

template <typename Kern>
int resident_slots(Kern kern, size_t lds, int (&cache)[16], int threads);

template <int P, typename TIN, class Plan>
int launch_rowgemm(const RowArgs& a, hipStream_t st, const char* what) {
  constexpr int TN = Plan::WPN * Plan::FB * 16, TP = Plan::WPP * Plan::PB * 16;
  if constexpr (P == PREC_BF16 && std::is_same<TIN, __bf16>::value && row_async_ok<Plan>()) {
    // bf16 operands: persistent workgroups on the asynchronous-copy ring,
    // one round of resident workgroups, each a contiguous range of tiles
    constexpr int TN = RowAsyncPol<Plan>::TN, TP = RowAsyncPol<Plan>::TP;
    const int64_t ntiles = ((a.npts + TP - 1) / TP) * ((a.nout + TN - 1) / TN);
    if (ntiles == 0) return NEB_OK;
    // the persistent walk needs one contraction length for every tile: a
    // block-diagonal window is used when the tile width and the block nest
    // and every window lies whole inside k (else the full k, same result)
    RowArgs b = a;
    if (b.kwin > 0) {
      const int64_t wl = TN >= b.kwin ? TN : b.kwin;
      if ((TN % b.kwin != 0 && b.kwin % TN != 0) || b.nout % wl != 0 || b.k < b.nout) b.kwin = 0;
    }
    auto kern = rowgemm_async_kernel<Plan>;
    constexpr int NTH = RowAsyncPol<Plan>::NTH;
    static int slots_cache[16] = {0};
    const int64_t slots = resident_slots(kern, 0, slots_cache, NTH);
    const int64_t per_wg = (ntiles + slots - 1) / slots;
    const int64_t grid = (ntiles + per_wg - 1) / per_wg;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(NTH), 0, st, b, per_wg);
    return hip_status(what);
  }
  const size_t lds = rowgemm_lds_bytes<P, Plan>();
  auto kern = rowgemm_kernel<P, TIN, Plan>;
  if (int rc = set_lds(kern, lds)) return rc;
  const int64_t gx = (a.npts + TP - 1) / TP, gy = (a.nout + TN - 1) / TN;
  if (gx * gy > 0x7fffffff) return fail(NEB_ERR_SHAPE, "%s: grid too large", what);
  if (gx == 0 || gy == 0) return NEB_OK;
  hipLaunchKernelGGL(kern, dim3((unsigned)(gx * gy)), dim3(256), lds, st, a);
  return hip_status(what);
}

// rowdiag_kernel launch: one thread per (point, group of GE features)
template <int P, typename TIN, class Plan>
int launch_rowdiag(const RowArgs& a, int64_t kdiag, hipStream_t st, const char* what) {
  constexpr int GE = Plan::template out_bytes<P>() == 2 ? 8 : 4;
  if (kdiag < GE || kdiag % GE) return fail(NEB_ERR_SHAPE, "%s: block must be a multiple of %d", what, GE);
  const int64_t tot = a.npts * ((a.nout + GE - 1) / GE);
  if (tot == 0) return NEB_OK;
  if ((tot + 255) / 256 > 0x7fffffff) return fail(NEB_ERR_SHAPE, "%s: grid too large", what);
  hipLaunchKernelGGL((rowdiag_kernel<P, TIN, Plan>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, a, kdiag);
  return hip_status(what);
}

template <int P> struct FieldT { using type = float; };
template <> struct FieldT<PREC_F64> { using type = double; };

RowArgs row_args() {
  RowArgs a;
  memset(&a, 0, sizeof(a));
  return a;
}

int check_common(int prec, int dim, int act) {
  if (prec != NEB_F32 && prec != NEB_BF16 && prec != NEB_F64) return fail(NEB_ERR_ARG, "unknown prec %d", prec);
  if (dim != 2 && dim != 3) return fail(NEB_ERR_ARG, "dim must be 2 or 3 (got %d)", dim);
  if (act != NEB_TANH && act != NEB_SIN) return fail(NEB_ERR_ARG, "unknown activation %d", act);
  return NEB_OK;
}

// The dim argument of the jet-layer entry points: 2, 3 (second-order jets)
// or NEB_DIM_2D_JET1 (2D first-order jets, kernel code JET1_2D).
int check_layer(int prec, int dim, int act) {
  return check_common(prec, dim == NEB_DIM_2D_JET1 ? 2 : dim, act);
}

bool is_stream_lift(int lift) { return lift == NEB_LIFT_2D_STREAM || lift == NEB_LIFT_2D_STREAM_JET1; }

// The lift argument of the output-layer entry points: 2 (2D velocity +
// vorticity), 3 (3D curl), NEB_LIFT_2D_HESSIAN (2D velocity + Hessian rows),
// NEB_LIFT_2D_STREAM / _JET1 (2D streamfunction + gradient rows, on second-
// or first-order tapes).
int check_lift(int prec, int lift, int act) {
  if (lift == NEB_LIFT_2D_HESSIAN || is_stream_lift(lift)) return check_common(prec, 2, act);
  return check_common(prec, lift, act);
}

// --------------------------------------------------------------------------
template <int P, int D, int ACT>
int layer_forward(int64_t n, int64_t k, int64_t width, const void* x, int64_t ldx, const void* w, int64_t ldw,
                  const void* bias, double omega, void* z, void* y, int64_t ldy, hipStream_t st, int64_t kwin = 0) {
  constexpr int C = Jet2Layout<D>::C;
  using Sx = typename Prec<P>::store_t;
  RowArgs a = row_args();
  for (int c = 0; c < C; ++c) {
    a.x[c] = reinterpret_cast<const Sx*>(x) + (int64_t)c * n * ldx;
    a.ldx[c] = ldx;
  }
  a.w[0] = w; a.ldw[0] = ldw;
  a.npts = n; a.nout = width; a.k = k;
  a.bias = bias; a.omega = omega;
  a.out0 = z; a.out1 = y; a.ldo = ldy; a.cstride = n * ldy;
  a.kwin = kwin;
  return launch_rowgemm<P, Sx, PlanHiddenFwd<D, ACT>>(a, st, "neb_jet_layer_forward");
}

template <int P, int D>
int output_jets(int64_t n, int64_t k, int64_t nout, const void* y, int64_t ldy, const void* w, int64_t ldw,
                const void* bias, void* out, int64_t ldo, hipStream_t st, const void* env = nullptr) {
  constexpr int C = Jet2Layout<D>::C;
  using Sx = typename Prec<P>::store_t;
  RowArgs a = row_args();
  for (int c = 0; c < C; ++c) {
    a.x[c] = reinterpret_cast<const Sx*>(y) + (int64_t)c * n * ldy;
    a.ldx[c] = ldy;
  }
  a.w[0] = w; a.ldw[0] = ldw;
  a.npts = n; a.nout = nout; a.k = k;
  a.bias = bias;
  a.out0 = out; a.ldo = ldo; a.cstride = n * ldo;
  if (env) {
    a.env = env;
    return launch_rowgemm<P, Sx, PlanOutJetsT<D, true>>(a, st, "neb_jet_output_jets_enveloped");
  }
  return launch_rowgemm<P, Sx, PlanOutJets<D>>(a, st, "neb_jet_output_jets");
}

template <typename SO, typename M, int D>
int envelope_reverse_launch(int64_t n, int64_t nout, const void* env, const void* adj, int64_t lda, void* out,
                            int64_t ldo, hipStream_t st) {
  const int64_t tot = n * nout;
  if (tot / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_envelope_reverse: too many elements");
  hipLaunchKernelGGL((envelope_reverse_kernel<SO, M, D>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const M*>(env), reinterpret_cast<const M*>(adj), lda, n, nout,
                     reinterpret_cast<SO*>(out), ldo);
  return hip_status("neb_envelope_reverse");
}

template <int P, bool SPLIT = false>
int output_fields(int dim, int64_t n, int64_t k, int64_t K, const void* y, int64_t ldy, const void* w, int64_t ldw,
                  void* mass, void* stiff, int64_t ldf, int layout, hipStream_t st, void* mass_lo = nullptr,
                  void* stiff_lo = nullptr) {
  using Sx = typename Prec<P>::store_t;
  const Sx* Y = reinterpret_cast<const Sx*>(y);
  const int64_t cs = n * ldy;
  RowArgs a = row_args();
  a.npts = n; a.nout = K; a.k = k;
  a.out0 = mass; a.out1 = stiff; a.ldo = ldf; a.ldo2 = ldf; a.layout = layout;
  a.out0lo = mass_lo; a.out1lo = stiff_lo;
  const char* what = SPLIT ? "neb_jet_output_fields_split" : "neb_jet_output_fields";
  if (is_stream_lift(dim)) {
    // channels v, gx, gy
    for (int c = 0; c < 3; ++c) { a.x[c] = Y + c * cs; a.ldx[c] = ldy; }
    a.w[0] = w; a.ldw[0] = ldw;
    return launch_rowgemm<P, Sx, PlanOutFields2ST<SPLIT, SPLIT>>(a, st, what);
  }
  if (dim == NEB_LIFT_2D_HESSIAN) {
    // channels gx, gy, hxx, hxy, hyy
    for (int c = 0; c < 5; ++c) { a.x[c] = Y + (1 + c) * cs; a.ldx[c] = ldy; }
    a.w[0] = w; a.ldw[0] = ldw;
    return launch_rowgemm<P, Sx, PlanOutFields2HT<SPLIT, SPLIT>>(a, st, what);
  }
  if (dim == 2) {
    // channels: 0 v, 1 gx, 2 gy, 3 hxx, 4 hxy, 5 hyy
    const int chans[4] = {1, 2, 3, 5};
    for (int c = 0; c < 4; ++c) { a.x[c] = Y + chans[c] * cs; a.ldx[c] = ldy; }
    a.w[0] = w; a.ldw[0] = ldw;
    return launch_rowgemm<P, Sx, PlanOutFields2T<SPLIT, SPLIT>>(a, st, what);
  }
  for (int c = 0; c < 9; ++c) { a.x[c] = Y + (1 + c) * cs; a.ldx[c] = ldy; }
  for (int s = 0; s < 3; ++s) {
    a.w[s] = reinterpret_cast<const Sx*>(w) + s * ldw;
    a.ldw[s] = 3 * ldw;
  }
  return launch_rowgemm<P, Sx, PlanOutFields3T<SPLIT>>(a, st, what);
}

// split operands: x_hi/x_lo, w_hi/w_lo planes (PlanHiddenFwdSplit)
template <int D, int ACT>
int layer_forward_x3(int64_t n, int64_t k, int64_t width, const void* x_hi, const void* x_lo, int64_t ldx,
                     const void* w_hi, const void* w_lo, int64_t ldw, const void* bias, double omega, void* z_hi,
                     void* z_lo, void* y_hi, void* y_lo, int64_t ldy, hipStream_t st, int64_t kwin = 0) {
  constexpr int C = Jet2Layout<D>::C;
  static_assert(2 * C <= RG_MAXCH, "split input planes");
  RowArgs a = row_args();
  for (int c = 0; c < C; ++c) {
    a.x[c] = reinterpret_cast<const __bf16*>(x_hi) + (int64_t)c * n * ldx;
    a.x[C + c] = reinterpret_cast<const __bf16*>(x_lo) + (int64_t)c * n * ldx;
    a.ldx[c] = a.ldx[C + c] = ldx;
  }
  a.w[0] = w_hi; a.w[1] = w_lo; a.ldw[0] = a.ldw[1] = ldw;
  a.npts = n; a.nout = width; a.k = k;
  a.bias = bias; a.omega = omega;
  a.out0 = z_hi; a.out0lo = z_lo; a.out1 = y_hi; a.out1lo = y_lo;
  a.ldo = ldy; a.cstride = n * ldy;
  a.kwin = kwin;
  return launch_rowgemm<PREC_BF16, __bf16, PlanHiddenFwdSplit<D, ACT>>(a, st, "neb_jet_layer_forward_x3");
}

int output_fields_x3(int dim, int64_t n, int64_t k, int64_t K, const void* y_hi, const void* y_lo, int64_t ldy,
                     const void* w_hi, const void* w_lo, int64_t ldw, void* mass_hi, void* mass_lo, void* stiff_hi,
                     void* stiff_lo, int64_t ldf, int layout, hipStream_t st) {
  const int64_t cs = n * ldy;
  const __bf16* Yh = reinterpret_cast<const __bf16*>(y_hi);
  const __bf16* Yl = reinterpret_cast<const __bf16*>(y_lo);
  RowArgs a = row_args();
  a.npts = n; a.nout = K; a.k = k;
  a.out0 = mass_hi; a.out0lo = mass_lo; a.out1 = stiff_hi; a.out1lo = stiff_lo;
  a.ldo = ldf; a.ldo2 = ldf; a.layout = layout;
  if (is_stream_lift(dim)) {
    for (int c = 0; c < 3; ++c) {
      a.x[c] = Yh + c * cs;
      a.x[3 + c] = Yl + c * cs;
      a.ldx[c] = a.ldx[3 + c] = ldy;
    }
    a.w[0] = w_hi; a.w[1] = w_lo; a.ldw[0] = a.ldw[1] = ldw;
    return launch_rowgemm<PREC_BF16, __bf16, PlanOutFields2SSplit>(a, st, "neb_jet_output_fields_x3");
  }
  if (dim == NEB_LIFT_2D_HESSIAN) {
    for (int c = 0; c < 5; ++c) {
      a.x[c] = Yh + (1 + c) * cs;
      a.x[5 + c] = Yl + (1 + c) * cs;
      a.ldx[c] = a.ldx[5 + c] = ldy;
    }
    a.w[0] = w_hi; a.w[1] = w_lo; a.ldw[0] = a.ldw[1] = ldw;
    return launch_rowgemm<PREC_BF16, __bf16, PlanOutFields2HSplit>(a, st, "neb_jet_output_fields_x3");
  }
  if (dim == 2) {
    const int chans[4] = {1, 2, 3, 5};  // gx, gy, hxx, hyy
    for (int c = 0; c < 4; ++c) {
      a.x[c] = Yh + chans[c] * cs;
      a.x[4 + c] = Yl + chans[c] * cs;
      a.ldx[c] = a.ldx[4 + c] = ldy;
    }
    a.w[0] = w_hi; a.w[1] = w_lo; a.ldw[0] = a.ldw[1] = ldw;
    return launch_rowgemm<PREC_BF16, __bf16, PlanOutFields2Split>(a, st, "neb_jet_output_fields_x3");
  }
  // 3D: channels g0..g2, h0..h5 (1..9); weight slices W_c = rows 3k+c
  for (int c = 0; c < 9; ++c) {
    a.x[c] = Yh + (1 + c) * cs;
    a.x[9 + c] = Yl + (1 + c) * cs;
    a.ldx[c] = a.ldx[9 + c] = ldy;
  }
  for (int sl = 0; sl < 3; ++sl) {
    a.w[sl] = reinterpret_cast<const __bf16*>(w_hi) + sl * ldw;
    a.w[3 + sl] = reinterpret_cast<const __bf16*>(w_lo) + sl * ldw;
    a.ldw[sl] = a.ldw[3 + sl] = 3 * ldw;
  }
  return launch_rowgemm<PREC_BF16, __bf16, PlanOutFields3Split>(a, st, "neb_jet_output_fields_x3");
}

// kdiag > 0: the weights are mode-diagonal, given as each row's one nonzero
// (bm / bs: [ncomp][ldb] vectors, entry f of component c the weight of
// output column f, which reads field column f / kdiag; launch_rowdiag);
// else a row GEMM over the K columns of [ncomp * width][ldb] weights.
template <int P, int ACT, typename F>
int output_backward(int dim, int64_t n, int64_t width, int64_t K, const void* mass, const void* stiff, int64_t ldf,
                    int layout, const void* bm, const void* bs, int64_t ldb, const void* z, int64_t ldz,
                    double omega, void* adj, int64_t lda, hipStream_t st, int64_t kdiag = 0) {
  using Sx = typename Prec<P>::store_t;
  const F* Mf = reinterpret_cast<const F*>(mass);
  const F* Sf = reinterpret_cast<const F*>(stiff);
  RowArgs a = row_args();
  a.npts = n; a.nout = width; a.k = K;
  a.tape = z; a.ldt = ldz; a.tstride = n * ldz;
  a.omega = omega;
  a.out0 = adj; a.ldo = lda; a.cstride = n * lda;
  if (is_stream_lift(dim)) {
    a.x[0] = Mf; a.ldx[0] = ldf;
    a.x[1] = Sf; a.ldx[1] = layout == 0 ? 2 * ldf : ldf;
    a.x[2] = layout == 0 ? Sf + ldf : Sf + n * ldf; a.ldx[2] = a.ldx[1];
    a.w[0] = bm; a.ldw[0] = ldb;
    a.w[1] = bs; a.ldw[1] = ldb;
    if (dim == NEB_LIFT_2D_STREAM_JET1)
      return kdiag > 0
                 ? launch_rowdiag<P, F, PlanOutBwd2S<ACT, JET1_2D>>(a, kdiag, st, "neb_jet_output_backward_blocks")
                 : launch_rowgemm<P, F, PlanOutBwd2S<ACT, JET1_2D>>(a, st, "neb_jet_output_backward");
    return kdiag > 0 ? launch_rowdiag<P, F, PlanOutBwd2S<ACT>>(a, kdiag, st, "neb_jet_output_backward_blocks")
                     : launch_rowgemm<P, F, PlanOutBwd2S<ACT>>(a, st, "neb_jet_output_backward");
  }
  if (dim == NEB_LIFT_2D_HESSIAN) {
    a.x[0] = Mf; a.ldx[0] = layout == 0 ? 2 * ldf : ldf;
    a.x[1] = layout == 0 ? Mf + ldf : Mf + n * ldf; a.ldx[1] = a.ldx[0];
    for (int r = 0; r < 3; ++r) {
      a.x[2 + r] = layout == 0 ? Sf + r * ldf : Sf + r * n * ldf;
      a.ldx[2 + r] = layout == 0 ? 3 * ldf : ldf;
    }
    a.w[0] = bm; a.ldw[0] = ldb;
    a.w[1] = bs; a.ldw[1] = ldb;
    return kdiag > 0 ? launch_rowdiag<P, F, PlanOutBwd2H<ACT>>(a, kdiag, st, "neb_jet_output_backward_blocks")
                     : launch_rowgemm<P, F, PlanOutBwd2H<ACT>>(a, st, "neb_jet_output_backward");
  }
  if (dim == 2) {
    if (layout == 0) {
      a.x[0] = Mf; a.ldx[0] = 2 * ldf;
      a.x[1] = Mf + ldf; a.ldx[1] = 2 * ldf;
    } else {
      a.x[0] = Mf; a.ldx[0] = ldf;
      a.x[1] = Mf + n * ldf; a.ldx[1] = ldf;
    }
    a.x[2] = Sf; a.ldx[2] = ldf;
    a.w[0] = bm; a.ldw[0] = ldb;
    a.w[1] = bs; a.ldw[1] = ldb;
    return kdiag > 0 ? launch_rowdiag<P, F, PlanOutBwd2<ACT>>(a, kdiag, st, "neb_jet_output_backward_blocks")
                     : launch_rowgemm<P, F, PlanOutBwd2<ACT>>(a, st, "neb_jet_output_backward");
  }
  for (int c = 0; c < 3; ++c) {
    a.x[c] = layout == 0 ? Mf + c * ldf : Mf + c * n * ldf;
    a.ldx[c] = layout == 0 ? 3 * ldf : ldf;
  }
  for (int c = 0; c < 9; ++c) {
    a.x[3 + c] = layout == 0 ? Sf + c * ldf : Sf + c * n * ldf;
    a.ldx[3 + c] = layout == 0 ? 9 * ldf : ldf;
  }
  const int64_t wstep = kdiag > 0 ? ldb : width * ldb;  // component s's weights
  for (int s = 0; s < 3; ++s) {
    a.w[s] = reinterpret_cast<const Sx*>(bm) + s * wstep;
    a.w[3 + s] = reinterpret_cast<const Sx*>(bs) + s * wstep;
    a.ldw[s] = ldb;
    a.ldw[3 + s] = ldb;
  }
  return kdiag > 0 ? launch_rowdiag<P, F, PlanOutBwd3<ACT>>(a, kdiag, st, "neb_jet_output_backward_blocks")
                     : launch_rowgemm<P, F, PlanOutBwd3<ACT>>(a, st, "neb_jet_output_backward");
}

template <int P, int D, int ACT>
int layer_backward(int64_t n, int64_t k, int64_t width, const void* adj_in, int64_t lda_in, const void* wt,
                   int64_t ldw, const void* z_prev, int64_t ldz, double omega_prev, void* adj_out,
                   int64_t lda_out, hipStream_t st, int64_t kwin = 0) {
  constexpr int C = Jet2Layout<D>::C;
  using Sx = typename Prec<P>::store_t;
  RowArgs a = row_args();
  for (int c = 0; c < C; ++c) {
    a.x[c] = reinterpret_cast<const Sx*>(adj_in) + (int64_t)c * n * lda_in;
    a.ldx[c] = lda_in;
  }
  a.w[0] = wt; a.ldw[0] = ldw;
  a.npts = n; a.nout = k; a.k = width;
  a.tape = z_prev; a.ldt = ldz; a.tstride = n * ldz;
  a.omega = omega_prev;
  a.out0 = adj_out; a.ldo = lda_out; a.cstride = n * lda_out;
  a.kwin = kwin;
  return launch_rowgemm<P, Sx, PlanHiddenBwd<D, ACT>>(a, st, "neb_jet_layer_backward");
}

// --------------------------------------------------------------------------
int cross_tiles(int64_t m, int64_t n, int symmetric, int* tm, int* tn) {
  *tm = (int)((m + CR_TM - 1) / CR_TM);
  *tn = (int)((n + CR_TN - 1) / CR_TN);
  return symmetric ? (*tm) * (*tm + 1) / 2 : (*tm) * (*tn);
}

// Tiles of a block-diagonal cross (m x m output, diagonal blocks of `block`):
// bdiag x bdiag tiles per block side (1 when a block fits inside a tile).
int cross_block_tiles(int64_t m, int64_t block, int* tm, int* bdiag) {
  *tm = (int)((m + CR_TM - 1) / CR_TM);
  *bdiag = block >= CR_TM ? (int)(block / CR_TM) : 1;
  return (*tm / *bdiag) * (*bdiag) * (*bdiag);
}

// Row splits of a cross product: as many as fill ONE round of resident
// workgroups (`slots` = CUs x resident workgroups per CU, at most 2 per CU,
// which bounds the FP64 partials), with ntiles * splits a multiple of 8 so
// every XCD gets the same number of work items, and >= 1024 rows per split
// -- or >= 256 where 1024-row splits would leave most CUs idle (a narrow
// product: a few output tiles over a few thousand rows); each split writes a
// partial tile, so more splits than the chip needs only add partial bytes.
// A second, partly filled round would idle most of the chip for as long as
// the first.
constexpr int CR_MAX_SLOTS = 512;

int cross_splits(int ntiles, int64_t rows, int slots = CR_MAX_SLOTS) {
  if (slots > CR_MAX_SLOTS) slots = CR_MAX_SLOTS;
  int s = slots / ntiles;
  while (s > 1 && (s * ntiles) % 8 != 0) --s;
  int64_t maxs = (rows + 1023) / 1024;
  if (maxs * ntiles < 256) maxs = (rows + 255) / 256;
  if (s > maxs) s = (int)maxs;
  if (s < 1) s = 1;
  return s;
}

// Resident workgroups that an asynchronous cross launch of the calling
// thread may use (0: as many as fit); see neb_set_cross_occupancy.
thread_local int g_cross_max_wg = 0;

// Resident workgroups of a kernel on the current device times the CU count
// (at most 2 per CU), queried once per kernel and device into `cache`.
template <typename Kern>
int resident_slots(Kern kern, size_t lds, int (&cache)[16], int threads) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return CR_MAX_SLOTS;
  if (cache[dev] == 0) {
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), threads, lds) !=
            hipSuccess ||
        cus < 1 || per_cu < 1) {
      (void)hipGetLastError();
      return CR_MAX_SLOTS;
    }
    cache[dev] = cus * (per_cu < 2 ? per_cu : 2);
  }
  return cache[dev];
}

template <int MODE, typename TX, typename TY>
int launch_cross(CrossArgs& a, double scale, double beta, double* G, int64_t ldg, hipStream_t st) {
  const size_t lds = cross_lds_bytes<MODE>();
  auto kern = cross_kernel<MODE, TX, TY>;
  if (int rc = set_lds(kern, lds)) return rc;
  // splits were sized for CR_MAX_SLOTS (the workspace bound); fit one round
  int64_t maxrows = 0;
  for (int s = 0; s < a.nseg; ++s) maxrows = a.seg[s].rows > maxrows ? a.seg[s].rows : maxrows;
  static int slots_cache[16] = {0};  // one per kernel instantiation
  const int fit = cross_splits(a.ntiles, maxrows, resident_slots(kern, lds, slots_cache, 256));
  if (fit < a.splits) a.splits = fit;
  hipLaunchKernelGGL(kern, dim3(a.ntiles * a.splits), dim3(256), lds, st, a);
  if (int rc = hip_status("neb_cross")) return rc;
  const int64_t nbi = (a.m + 31) / 32, nbj = (a.n + 31) / 32;
  const int64_t nblk = a.symmetric ? nbi * (nbi + 1) / 2 : nbi * nbj;
  hipLaunchKernelGGL(cross_reduce_kernel<double>, dim3((unsigned)nblk), dim3(256), 0, st, a.partial, a.splits, a.ntiles,
                     a.tiles_n, a.symmetric, a.bdiag, a.m, a.n, scale, beta, G, ldg);
  return hip_status("neb_cross reduce");
}

// bf16-stored operands (one plane, or hi + lo planes) through the
// asynchronous global -> LDS pipeline (cross_async.hpp)
// ring / wave configuration per plane count (measured on the K = 1024
// workloads): one plane: 8 waves of 32 x 64, 4 slots (3 k-steps of copies in
// flight), two workgroups per CU; two planes: 4 waves of 64 x 64 (fewer
// fragment reads per MFMA), 2 slots, two workgroups per CU (160 KB LDS).
template <int NL> struct AsyncCfg;
template <> struct AsyncCfg<1> { static constexpr int WM = 4, WN = 2, NS = 4; };
template <> struct AsyncCfg<2> { static constexpr int WM = 2, WN = 2, NS = 2; };

template <int NL, bool NEG>
int launch_cross_async_as(CrossArgs& a, double scale, double beta, double* G, int64_t ldg, hipStream_t st) {
  using C = AsyncCfg<NL>;
  constexpr int NT = 64 * C::WM * C::WN;
  const size_t lds = 0;  // the ring is static LDS
  auto kern = cross_async_kernel<NL, C::WM, C::WN, C::NS, NEG>;
  int64_t maxrows = 0;
  for (int s = 0; s < a.nseg; ++s) maxrows = a.seg[s].rows > maxrows ? a.seg[s].rows : maxrows;
  static int slots_cache[16] = {0};  // one per kernel instantiation
  int slots = resident_slots(kern, lds, slots_cache, NT);
  if (g_cross_max_wg > 0 && slots > g_cross_max_wg) slots = g_cross_max_wg;
  const int fit = cross_splits(a.ntiles, maxrows, slots);
  if (fit < a.splits) a.splits = fit;
  hipLaunchKernelGGL(kern, dim3(a.ntiles * a.splits), dim3(NT), lds, st, a);
  if (int rc = hip_status("neb_cross (async)")) return rc;
  const int64_t nbi = (a.m + 31) / 32, nbj = (a.n + 31) / 32;
  const int64_t nblk = a.symmetric ? nbi * (nbi + 1) / 2 : nbi * nbj;
  hipLaunchKernelGGL(cross_reduce_kernel<float>, dim3((unsigned)nblk), dim3(256), 0, st,
                     reinterpret_cast<const float*>(a.partial), a.splits, a.ntiles,
                     a.tiles_n, a.symmetric, a.bdiag, a.m, a.n, scale, beta, G, ldg);
  return hip_status("neb_cross reduce");
}

// sign flips compiled in only when some segment has sign -1
template <int NL>
int launch_cross_async(CrossArgs& a, double scale, double beta, double* G, int64_t ldg, hipStream_t st) {
  bool neg = false;
  for (int s = 0; s < a.nseg; ++s) neg = neg || a.seg[s].sign < 0;
  return neg ? launch_cross_async_as<NL, true>(a, scale, beta, G, ldg, st)
             : launch_cross_async_as<NL, false>(a, scale, beta, G, ldg, st);
}

// The asynchronous path takes 16-byte aligned bf16 rows, no row weights,
// segment signs of +-1 and row strides below 2^25 elements (the per-lane
// copy offsets of a 32-row k-step are 32-bit byte offsets).
bool async_ok(const CrossArgs& a) {
  if (a.roww) return false;
  for (int s = 0; s < a.nseg; ++s) {
    const CrossSeg& g = a.seg[s];
    if (g.sign != 1.0 && g.sign != -1.0) return false;
    if (g.ldx >= ((int64_t)1 << 25) || g.ldy >= ((int64_t)1 << 25)) return false;
    if (!aligned(g.x, g.ldx, NEB_BF16) || !aligned(g.y, g.ldy, NEB_BF16)) return false;
    if ((g.xlo && !aligned(g.xlo, g.ldx, NEB_BF16)) || (g.ylo && !aligned(g.ylo, g.ldy, NEB_BF16))) return false;
  }
  return true;
}

template <int MODE>
int cross_dispatch_types(int xd, int yd, CrossArgs& a, double scale, double beta, double* G, int64_t ldg,
                         hipStream_t st) {
  if constexpr (MODE == CR_F64) {
    if (xd == NEB_F64 && yd == NEB_F64) return launch_cross<MODE, double, double>(a, scale, beta, G, ldg, st);
    // fp32 operands (network fields) widened exactly to fp64 on load: the
    // FP64 mass and projection loads of an fp32 basis
    if (xd == NEB_F32 && yd == NEB_F64) return launch_cross<MODE, float, double>(a, scale, beta, G, ldg, st);
    if (xd == NEB_F32 && yd == NEB_F32) return launch_cross<MODE, float, float>(a, scale, beta, G, ldg, st);
    return fail(NEB_ERR_ARG, "NEB_CROSS_F64 needs fp64 or fp32 X with fp64 / fp32 Y");
  } else if constexpr (MODE == CR_F32) {
    if (xd == NEB_F32 && yd == NEB_F32) return launch_cross<MODE, float, float>(a, scale, beta, G, ldg, st);
    return fail(NEB_ERR_ARG, "NEB_CROSS_F32 needs fp32 operands");
  } else {
    if (xd == NEB_F32 && yd == NEB_F32) return launch_cross<MODE, float, float>(a, scale, beta, G, ldg, st);
    if (MODE == CR_BF16 && xd == NEB_BF16 && yd == NEB_BF16 && async_ok(a))
      return launch_cross_async<1>(a, scale, beta, G, ldg, st);
    if (xd == NEB_BF16 && yd == NEB_BF16) return launch_cross<MODE, __bf16, __bf16>(a, scale, beta, G, ldg, st);
    if (xd == NEB_F32 && yd == NEB_BF16) return launch_cross<MODE, float, __bf16>(a, scale, beta, G, ldg, st);
    if (xd == NEB_BF16 && yd == NEB_F32) return launch_cross<MODE, __bf16, float>(a, scale, beta, G, ldg, st);
    return fail(NEB_ERR_ARG, "bf16 cross modes need fp32/bf16 operands");
  }
}

template <typename T> int colsum_launch(const void* x, int64_t rows, int64_t cols, int64_t ld, double beta,
                                        double* out, double* part, int splits, hipStream_t st) {
  const int64_t chunk = (rows + splits - 1) / splits;
  const int64_t ngroups = (cols + 16 / (int64_t)sizeof(T) - 1) / (16 / (int64_t)sizeof(T));
  const int64_t slab = colsum_slab_groups<T>(cols), slabs = (ngroups + slab - 1) / slab;
  if (slabs > 65535) return fail(NEB_ERR_SHAPE, "neb_colsum: too many columns");
  hipLaunchKernelGGL(colsum_kernel<T>, dim3(splits, (unsigned)slabs), dim3(256), 0, st, reinterpret_cast<const T*>(x),
                     ld, rows, cols, chunk, part);
  if (int rc = hip_status("neb_colsum")) return rc;
  hipLaunchKernelGGL(sum_splits_kernel, sum_splits_grid(cols), dim3(256), 0, st, part, splits, cols,
                     out, beta);
  return hip_status("neb_colsum reduce");
}

int colsum_splits(int64_t rows) {
  int64_t s = (rows + 127) / 128;
  if (s > 1024) s = 1024;
  if (s < 1) s = 1;
  return (int)s;
}

template <typename TS, typename TD>
int stage_launch(const void* src, int64_t rows, int64_t cols, int64_t lds, void* dst, int64_t ldd, int64_t drows,
                 int64_t dcols, int transpose, double scale, hipStream_t st) {
  const int64_t tot = drows * dcols;
  if (tot == 0) return NEB_OK;
  hipLaunchKernelGGL((stage_kernel<TS, TD>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const TS*>(src), rows, cols, lds, reinterpret_cast<TD*>(dst), ldd, drows, dcols,
                     transpose, scale);
  return hip_status("neb_stage");
}

template <typename TS>
int stage_dst(int dst_dtype, const void* src, int64_t rows, int64_t cols, int64_t lds, void* dst, int64_t ldd,
              int64_t drows, int64_t dcols, int transpose, double scale, hipStream_t st) {
  switch (dst_dtype) {
    case NEB_F32: return stage_launch<TS, float>(src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
    case NEB_BF16: return stage_launch<TS, __bf16>(src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
    case NEB_F64: return stage_launch<TS, double>(src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
  }
  return fail(NEB_ERR_ARG, "neb_stage: unknown dst dtype %d", dst_dtype);
}

// Split Gram on caller-provided partial-tile memory (the workspace, or the
// library's own scratch for neb_gram_split_noalloc); `cap` bytes available,
// `elem` bytes per partial value counted against it.
int gram_split_launch(const char* what, const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld,
                      double w, double beta, double* G, int64_t ldg, void* partial, size_t cap, size_t elem,
                      void* stream) {
  CrossArgs a;
  memset(&a, 0, sizeof(a));
  a.seg[0].x = phi_hi; a.seg[0].y = phi_hi;
  a.seg[0].xlo = phi_lo; a.seg[0].ylo = phi_lo;
  a.seg[0].ldx = ld; a.seg[0].ldy = ld;
  a.seg[0].rows = rows; a.seg[0].sign = 1.0;
  a.nseg = 1; a.m = K; a.n = K; a.roww = nullptr; a.symmetric = 1;
  a.ntiles = cross_tiles(K, K, 1, &a.tiles_m, &a.tiles_n);
  a.splits = cross_splits(a.ntiles, rows);
  const size_t need = (size_t)a.splits * a.ntiles * CR_TM * CR_TN * elem;
  if (cap < need) return fail(NEB_ERR_ARG, "%s: %zu bytes of partial tiles needed, %zu available", what, need, cap);
  a.partial = reinterpret_cast<double*>(partial);
  return launch_cross_async<2>(a, w, beta, G, ldg, S(stream));
}

int gram_split_check(const char* what, const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld,
                     const double* G, int64_t ldg) {
  if (rows < 0 || K < 1 || ld < K || ld >= (1 << 25) || ldg < K)
    return fail(NEB_ERR_SHAPE, "%s: bad sizes (ld must be below 2^25)", what);
  if ((rows > 0 && (!phi_hi || !phi_lo)) || !G) return fail(NEB_ERR_ARG, "%s: null pointer", what);
  if (!aligned(phi_hi, ld, NEB_BF16) || !aligned(phi_lo, ld, NEB_BF16))
    return fail(NEB_ERR_ARG, "%s: rows must be 16-byte aligned", what);
  return NEB_OK;
}

}  // namespace

// Scratch of neb_gram_split_noalloc: the fp32 split-R partial tiles.  Each
// GPU the library runs on holds its own copy from the module's load until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the
// host the name is a stand-in, not a device address): it uses the per-device
// address from hipGetSymbolAddress below.
static __device__ __align__(256) unsigned char scratch[NEB_NOALLOC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(NEB_ERR_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= 64) return fail(NEB_ERR_HIP, "device index %d outside 0..63", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NEB_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NEB_OK;
}

// Row chunks of neb_cross_modes: about 1,024 workgroups over the chunks and
// the column groups, at least 8 rows each.  Returns (chunks, rows per chunk).
std::pair<int64_t, int64_t> mode_cross_chunks(int64_t B, int64_t h, int64_t rows) {
  const int64_t gy = (B * h / 8 + 255) / 256;
  const int64_t want = gy >= 1024 ? 1 : 1024 / gy;
  int64_t chunk = (rows + want - 1) / want;
  if (chunk < 8) chunk = 8;
  const int64_t n = rows > 0 ? (rows + chunk - 1) / chunk : 1;
  return {n, chunk};
}

}  // namespace

// ==========================================================================
extern "C" {

size_t neb_cross_modes_workspace_size(int64_t B, int64_t h, int64_t rows) {
  return (size_t)mode_cross_chunks(B, h, rows).first * (size_t)(B * h) * sizeof(double);
}

int neb_cross_modes(int nseg, const void* const* x, const void* const* y, const int64_t* ldx, const int64_t* ldy,
                    const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t B, int64_t h,
                    double scale, double beta, double* out, void* workspace, size_t ws_bytes, void* stream) {
  if (nseg < 1 || nseg > MC_MAXSEG) return fail(NEB_ERR_ARG, "neb_cross_modes: 1..%d segments", MC_MAXSEG);
  if (B < 1 || h < 8 || h % 8) return fail(NEB_ERR_SHAPE, "neb_cross_modes: h must be a positive multiple of 8");
  if (!x || !y || !ldx || !ldy || !rows || !sign || !out || !workspace)
    return fail(NEB_ERR_ARG, "neb_cross_modes: null pointer");
  const bool xok = x_dtype == NEB_F32 || x_dtype == NEB_F64;
  const bool yok = y_dtype == NEB_BF16 || y_dtype == NEB_F32 || y_dtype == NEB_F64;
  if (!xok || !yok) return fail(NEB_ERR_ARG, "neb_cross_modes: x dtype F32 / F64, y dtype BF16 / F32 / F64");
  ModeCrossArgs a;
  memset(&a, 0, sizeof(a));
  int64_t rmax = 0;
  for (int s = 0; s < nseg; ++s) {
    if (rows[s] < 0 || ldx[s] < B || ldy[s] < B * h) return fail(NEB_ERR_SHAPE, "neb_cross_modes: segment %d sizes", s);
    if (rows[s] > 0 && (!x[s] || !y[s] || !aligned(y[s], ldy[s], y_dtype)))
      return fail(NEB_ERR_ARG, "neb_cross_modes: segment %d: null or unaligned (16-byte rows of y)", s);
    a.x[s] = x[s]; a.y[s] = y[s]; a.ldx[s] = ldx[s]; a.ldy[s] = ldy[s]; a.rows[s] = rows[s]; a.sign[s] = sign[s];
    rmax = rows[s] > rmax ? rows[s] : rmax;
  }
  const auto ch = mode_cross_chunks(B, h, rmax);
  if (ws_bytes < (size_t)ch.first * (size_t)(B * h) * sizeof(double))
    return fail(NEB_ERR_ARG, "neb_cross_modes: workspace too small");
  a.nseg = nseg; a.B = B; a.h = h; a.chunk = ch.second; a.scale = scale;
  a.partial = reinterpret_cast<double*>(workspace);
  hipStream_t st = S(stream);
  const int64_t gy = (B * h / 8 + 255) / 256;
  if (gy > 65535) return fail(NEB_ERR_SHAPE, "neb_cross_modes: too many columns");
  const dim3 grid((unsigned)ch.first, (unsigned)gy);
#define MCK(TX, TY) hipLaunchKernelGGL((mode_cross_kernel<TX, TY>), grid, dim3(256), 0, st, a)
  if (x_dtype == NEB_F32) {
    if (y_dtype == NEB_BF16) MCK(float, __bf16);
    else if (y_dtype == NEB_F32) MCK(float, float);
    else MCK(float, double);
  } else {
    if (y_dtype == NEB_BF16) MCK(double, __bf16);
    else if (y_dtype == NEB_F32) MCK(double, float);
    else MCK(double, double);
  }
#undef MCK
  if (int rc = hip_status("neb_cross_modes")) return rc;
  hipLaunchKernelGGL(sum_splits_kernel, sum_splits_grid(B * h), dim3(256), 0, st, a.partial,
                     (int)ch.first, B * h, out, beta);
  return hip_status("neb_cross_modes reduce");
}

int neb_set_cross_occupancy(int max_workgroups) {
  if (max_workgroups < 0) return fail(NEB_ERR_ARG, "neb_set_cross_occupancy: max_workgroups must be >= 0");
  g_cross_max_wg = max_workgroups;
  return NEB_OK;
}

int neb_abi_version(void) { return NEB_ABI_VERSION; }
const char* neb_last_error(void) { return g_err.c_str(); }

int neb_feature_jets(int prec, int dim, int64_t n, const void* x, int freqs, const double* lo, const double* hi,
                     void* out, int64_t ldf, void* stream) {
  if (int rc = check_common(prec, dim, NEB_TANH)) return rc;
  if (n < 0 || freqs < 1 || ldf < 2 * freqs * dim) return fail(NEB_ERR_SHAPE, "neb_feature_jets: bad sizes");
  if (n == 0) return NEB_OK;  // no points: nothing to read or write (buffers may be null)
  if (!x || !out || !lo || !hi) return fail(NEB_ERR_ARG, "neb_feature_jets: null pointer");
  const int64_t tot = n * ldf;
  const dim3 grid((unsigned)((tot + 255) / 256));
  const double l2 = dim == 3 ? lo[2] : 0.0, h2 = dim == 3 ? hi[2] : 1.0;
#define FEAT(D, TX, SS)                                                                                          \
  hipLaunchKernelGGL((feature_kernel<D, TX, SS>), grid, dim3(256), 0, S(stream), reinterpret_cast<const TX*>(x), \
                     n, freqs, lo[0], lo[1], l2, hi[0], hi[1], h2, reinterpret_cast<SS*>(out), ldf)
  if (prec == NEB_F64) {
    if (dim == 2) FEAT(2, double, double); else FEAT(3, double, double);
  } else if (prec == NEB_F32) {
    if (dim == 2) FEAT(2, float, float); else FEAT(3, float, float);
  } else {
    if (dim == 2) FEAT(2, float, __bf16); else FEAT(3, float, __bf16);
  }
#undef FEAT
  return hip_status("neb_feature_jets");
}

#define DISPATCH_PDA(CALL)                                                        \
  if (dim == NEB_DIM_2D_JET1) {                                                   \
    if (prec == NEB_BF16) return act == NEB_TANH ? CALL(PREC_BF16, JET1_2D, ACT_TANH) : CALL(PREC_BF16, JET1_2D, ACT_SIN); \
    if (prec == NEB_F32) return act == NEB_TANH ? CALL(PREC_F32, JET1_2D, ACT_TANH) : CALL(PREC_F32, JET1_2D, ACT_SIN);    \
    return act == NEB_TANH ? CALL(PREC_F64, JET1_2D, ACT_TANH) : CALL(PREC_F64, JET1_2D, ACT_SIN);                         \
  }                                                                               \
  switch (prec) {                                                                 \
    case NEB_BF16:                                                                \
      if (dim == 2) { if (act == NEB_TANH) return CALL(PREC_BF16, 2, ACT_TANH); return CALL(PREC_BF16, 2, ACT_SIN); } \
      if (act == NEB_TANH) return CALL(PREC_BF16, 3, ACT_TANH); return CALL(PREC_BF16, 3, ACT_SIN);                  \
    case NEB_F32:                                                                 \
      if (dim == 2) { if (act == NEB_TANH) return CALL(PREC_F32, 2, ACT_TANH); return CALL(PREC_F32, 2, ACT_SIN); }   \
      if (act == NEB_TANH) return CALL(PREC_F32, 3, ACT_TANH); return CALL(PREC_F32, 3, ACT_SIN);                    \
    default:                                                                      \
      if (dim == 2) { if (act == NEB_TANH) return CALL(PREC_F64, 2, ACT_TANH); return CALL(PREC_F64, 2, ACT_SIN); }   \
      if (act == NEB_TANH) return CALL(PREC_F64, 3, ACT_TANH); return CALL(PREC_F64, 3, ACT_SIN);                    \
  }

int neb_jet_layer_forward(int prec, int dim, int act, int64_t n, int64_t k, int64_t width, const void* x,
                          int64_t ldx, const void* w, int64_t ldw, const void* bias, double omega, void* z, void* y,
                          int64_t ldy, void* stream) {
  if (int rc = check_layer(prec, dim, act)) return rc;
  if (n < 0 || k < 1 || width < 1 || ldx < k || ldw < k || ldy < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_forward: bad sizes");
  if (n == 0) return NEB_OK;
  if (!x || !w || !z || !y) return fail(NEB_ERR_ARG, "neb_jet_layer_forward: null pointer");
  if (!aligned(x, ldx, prec) || !aligned(w, ldw, prec) || !aligned(y, ldy, prec) || !aligned(z, ldy, prec))
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward: rows must be 16-byte aligned");
#define CALL(P, D, A) layer_forward<P, D, A>(n, k, width, x, ldx, w, ldw, bias, omega, z, y, ldy, S(stream))
  DISPATCH_PDA(CALL)
#undef CALL
}

int neb_jet_layer_forward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block, const void* x,
                                 int64_t ldx, const void* w, int64_t ldw, const void* bias, double omega, void* z,
                                 void* y, int64_t ldy, void* stream) {
  if (int rc = check_layer(prec, dim, act)) return rc;
  if (n < 0 || width < 1 || block < 8 || block % 8 || width % block || ldx < width || ldw < width || ldy < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_forward_blocks: bad sizes (block a multiple of 8 dividing the width)");
  if (n == 0) return NEB_OK;
  if (!x || !w || !z || !y) return fail(NEB_ERR_ARG, "neb_jet_layer_forward_blocks: null pointer");
  if (!aligned(x, ldx, prec) || !aligned(w, ldw, prec) || !aligned(y, ldy, prec) || !aligned(z, ldy, prec))
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward_blocks: rows must be 16-byte aligned");
#define CALL(P, D, A) layer_forward<P, D, A>(n, width, width, x, ldx, w, ldw, bias, omega, z, y, ldy, S(stream), block)
  DISPATCH_PDA(CALL)
#undef CALL
}

int neb_jet_layer_backward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block,
                                  const void* adj_in, int64_t lda_in, const void* wt, int64_t ldw, const void* z_prev,
                                  int64_t ldz, double omega_prev, void* adj_out, int64_t lda_out, void* stream) {
  if (int rc = check_layer(prec, dim, act)) return rc;
  if (n < 0 || width < 1 || block < 8 || block % 8 || width % block || lda_in < width || ldw < width ||
      ldz < width || lda_out < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_backward_blocks: bad sizes (block a multiple of 8 dividing the width)");
  if (n == 0) return NEB_OK;
  if (!adj_in || !wt || !z_prev || !adj_out) return fail(NEB_ERR_ARG, "neb_jet_layer_backward_blocks: null pointer");
  if (!aligned(adj_in, lda_in, prec) || !aligned(wt, ldw, prec) || !aligned(z_prev, ldz, prec) ||
      !aligned(adj_out, lda_out, prec))
    return fail(NEB_ERR_ARG, "neb_jet_layer_backward_blocks: rows must be 16-byte aligned");
#define CALL(P, D, A)                                                                                            \
  layer_backward<P, D, A>(n, width, width, adj_in, lda_in, wt, ldw, z_prev, ldz, omega_prev, adj_out, lda_out, \
                          S(stream), block)
  DISPATCH_PDA(CALL)
#undef CALL
}

int neb_jet_output_jets_enveloped(int prec, int dim, int64_t n, int64_t k, int64_t nout, const void* y,
                                  int64_t ldy, const void* w, int64_t ldw, const void* bias, const void* env,
                                  void* out, int64_t ldo, void* stream) {
  if (int rc = check_common(prec, dim, NEB_TANH)) return rc;
  if (n < 0 || k < 1 || nout < 1 || ldy < k || ldw < k || ldo < nout)
    return fail(NEB_ERR_SHAPE, "neb_jet_output_jets_enveloped: bad sizes");
  if (n == 0) return NEB_OK;
  if (!y || !w || !out || !env) return fail(NEB_ERR_ARG, "neb_jet_output_jets_enveloped: null pointer");
  if (!aligned(y, ldy, prec) || !aligned(w, ldw, prec))
    return fail(NEB_ERR_ARG, "neb_jet_output_jets_enveloped: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
  if (prec == NEB_BF16)
    return dim == 2 ? output_jets<PREC_BF16, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env)
                    : output_jets<PREC_BF16, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env);
  if (prec == NEB_F32)
    return dim == 2 ? output_jets<PREC_F32, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env)
                    : output_jets<PREC_F32, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env);
  return dim == 2 ? output_jets<PREC_F64, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env)
                  : output_jets<PREC_F64, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st, env);
}

int neb_envelope_reverse(int prec, int dim, int64_t n, int64_t nout, const void* env, const void* adj, int64_t lda,
                         void* out, int64_t ldo, void* stream) {
  if (int rc = check_common(prec, dim, NEB_TANH)) return rc;
  if (n < 0 || nout < 1 || lda < nout || ldo < nout) return fail(NEB_ERR_SHAPE, "neb_envelope_reverse: bad sizes");
  if (n == 0) return NEB_OK;
  if (!env || !adj || !out) return fail(NEB_ERR_ARG, "neb_envelope_reverse: null pointer");
  hipStream_t st = S(stream);
  if (prec == NEB_BF16)
    return dim == 2 ? envelope_reverse_launch<__bf16, float, 2>(n, nout, env, adj, lda, out, ldo, st)
                    : envelope_reverse_launch<__bf16, float, 3>(n, nout, env, adj, lda, out, ldo, st);
  if (prec == NEB_F32)
    return dim == 2 ? envelope_reverse_launch<float, float, 2>(n, nout, env, adj, lda, out, ldo, st)
                    : envelope_reverse_launch<float, float, 3>(n, nout, env, adj, lda, out, ldo, st);
  return dim == 2 ? envelope_reverse_launch<double, double, 2>(n, nout, env, adj, lda, out, ldo, st)
                  : envelope_reverse_launch<double, double, 3>(n, nout, env, adj, lda, out, ldo, st);
}

int neb_jet_output_jets(int prec, int dim, int64_t n, int64_t k, int64_t nout, const void* y, int64_t ldy,
                        const void* w, int64_t ldw, const void* bias, void* out, int64_t ldo, void* stream) {
  if (int rc = check_common(prec, dim, NEB_TANH)) return rc;
  if (n < 0 || k < 1 || nout < 1 || ldy < k || ldw < k || ldo < nout)
    return fail(NEB_ERR_SHAPE, "neb_jet_output_jets: bad sizes");
  if (n == 0) return NEB_OK;
  if (!y || !w || !out) return fail(NEB_ERR_ARG, "neb_jet_output_jets: null pointer");
  if (!aligned(y, ldy, prec) || !aligned(w, ldw, prec))
    return fail(NEB_ERR_ARG, "neb_jet_output_jets: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
  if (prec == NEB_BF16) return dim == 2 ? output_jets<PREC_BF16, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st)
                                        : output_jets<PREC_BF16, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st);
  if (prec == NEB_F32) return dim == 2 ? output_jets<PREC_F32, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st)
                                       : output_jets<PREC_F32, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st);
  return dim == 2 ? output_jets<PREC_F64, 2>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st)
                  : output_jets<PREC_F64, 3>(n, k, nout, y, ldy, w, ldw, bias, out, ldo, st);
}

int neb_jet_output_fields(int prec, int dim, int64_t n, int64_t k, int64_t K, const void* y, int64_t ldy,
                          const void* w, int64_t ldw, void* mass, void* stiff, int64_t ldf, int layout,
                          void* stream) {
  if (int rc = check_lift(prec, dim, NEB_TANH)) return rc;
  if (n < 0 || k < 1 || K < 1 || ldy < k || ldw < k || ldf < K || (layout != 0 && layout != 1))
    return fail(NEB_ERR_SHAPE, "neb_jet_output_fields: bad sizes");
  if (n == 0) return NEB_OK;
  if (!y || !w || !mass || !stiff) return fail(NEB_ERR_ARG, "neb_jet_output_fields: null pointer");
  const int fdt = prec == NEB_F64 ? NEB_F64 : NEB_F32;
  if (!aligned(y, ldy, prec) || !aligned(w, ldw, prec) || !aligned(mass, ldf, fdt) || !aligned(stiff, ldf, fdt))
    return fail(NEB_ERR_ARG, "neb_jet_output_fields: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
  if (prec == NEB_BF16) return output_fields<PREC_BF16>(dim, n, k, K, y, ldy, w, ldw, mass, stiff, ldf, layout, st);
  if (prec == NEB_F32) return output_fields<PREC_F32>(dim, n, k, K, y, ldy, w, ldw, mass, stiff, ldf, layout, st);
  return output_fields<PREC_F64>(dim, n, k, K, y, ldy, w, ldw, mass, stiff, ldf, layout, st);
}

int neb_jet_output_fields_split(int dim, int64_t n, int64_t k, int64_t K, const void* y, int64_t ldy, const void* w,
                                int64_t ldw, void* mass_hi, void* mass_lo, void* stiff_hi, void* stiff_lo,
                                int64_t ldf, int layout, void* stream) {
  if (int rc = check_lift(NEB_BF16, dim, NEB_TANH)) return rc;
  if (n < 0 || k < 1 || K < 1 || ldy < k || ldw < k || ldf < K || (layout != 0 && layout != 1))
    return fail(NEB_ERR_SHAPE, "neb_jet_output_fields_split: bad sizes");
  if (n == 0) return NEB_OK;
  if (!y || !w || !mass_hi || !mass_lo || !stiff_hi || !stiff_lo)
    return fail(NEB_ERR_ARG, "neb_jet_output_fields_split: null pointer");
  if (!aligned(y, ldy, NEB_BF16) || !aligned(w, ldw, NEB_BF16) || !aligned(mass_hi, ldf, NEB_BF16) ||
      !aligned(mass_lo, ldf, NEB_BF16) || !aligned(stiff_hi, ldf, NEB_BF16) || !aligned(stiff_lo, ldf, NEB_BF16))
    return fail(NEB_ERR_ARG, "neb_jet_output_fields_split: rows must be 16-byte aligned");
  return output_fields<PREC_BF16, true>(dim, n, k, K, y, ldy, w, ldw, mass_hi, stiff_hi, ldf, layout, S(stream),
                                        mass_lo, stiff_lo);
}

static int output_backward_entry(int prec, int dim, int act, int64_t n, int64_t width, int64_t K, const void* mass,
                                 const void* stiff, int64_t ldf, int fields_dtype, int layout, const void* bm,
                                 const void* bs, int64_t ldb, const void* z, int64_t ldz, double omega, void* adj,
                                 int64_t lda, void* stream, int64_t kdiag);

int neb_jet_output_backward(int prec, int dim, int act, int64_t n, int64_t width, int64_t K, const void* mass,
                            const void* stiff, int64_t ldf, int fields_dtype, int layout, const void* bm,
                            const void* bs, int64_t ldb, const void* z, int64_t ldz, double omega, void* adj,
                            int64_t lda, void* stream) {
  return output_backward_entry(prec, dim, act, n, width, K, mass, stiff, ldf, fields_dtype, layout, bm, bs, ldb, z,
                               ldz, omega, adj, lda, stream, 0);
}

int neb_jet_output_backward_blocks(int prec, int dim, int act, int64_t n, int64_t width, int64_t block,
                                   const void* mass, const void* stiff, int64_t ldf, int fields_dtype, int layout,
                                   const void* bm, const void* bs, int64_t ldb, const void* z, int64_t ldz,
                                   double omega, void* adj, int64_t lda, void* stream) {
  if (block < 8 || block % 8 || width % block)
    return fail(NEB_ERR_SHAPE, "neb_jet_output_backward_blocks: block must be a multiple of 8 dividing the width");
  if (ldb < width) return fail(NEB_ERR_SHAPE, "neb_jet_output_backward_blocks: weight vectors shorter than the width");
  return output_backward_entry(prec, dim, act, n, width, width / block, mass, stiff, ldf, fields_dtype, layout, bm, bs,
                               ldb, z, ldz, omega, adj, lda, stream, block);
}

static int output_backward_entry(int prec, int dim, int act, int64_t n, int64_t width, int64_t K, const void* mass,
                                 const void* stiff, int64_t ldf, int fields_dtype, int layout, const void* bm,
                                 const void* bs, int64_t ldb, const void* z, int64_t ldz, double omega, void* adj,
                                 int64_t lda, void* stream, int64_t kdiag) {
  if (int rc = check_lift(prec, dim, act)) return rc;
  if (n < 0 || width < 1 || K < 1 || ldf < K || ldb < K || ldz < width || lda < width ||
      (layout != 0 && layout != 1))
    return fail(NEB_ERR_SHAPE, "neb_jet_output_backward: bad sizes");
  if (n == 0) return NEB_OK;
  if (!mass || !stiff || !bm || !bs || !z || !adj) return fail(NEB_ERR_ARG, "neb_jet_output_backward: null pointer");
  const bool fok = prec == NEB_F64 ? fields_dtype == NEB_F64
                                   : (fields_dtype == NEB_F32 || (prec == NEB_BF16 && fields_dtype == NEB_BF16));
  if (!fok) return fail(NEB_ERR_ARG, "neb_jet_output_backward: fields dtype %d does not fit prec %d", fields_dtype, prec);
  if (!aligned(mass, ldf, fields_dtype) || !aligned(stiff, ldf, fields_dtype) || !aligned(bm, ldb, prec) ||
      !aligned(bs, ldb, prec) || !aligned(z, ldz, prec) || !aligned(adj, lda, prec))
    return fail(NEB_ERR_ARG, "neb_jet_output_backward: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
#define OB(P, A, F) \
  output_backward<P, A, F>(dim, n, width, K, mass, stiff, ldf, layout, bm, bs, ldb, z, ldz, omega, adj, lda, st, kdiag)
  if (prec == NEB_BF16) {
    if (fields_dtype == NEB_BF16)
      return act == NEB_TANH ? OB(PREC_BF16, ACT_TANH, __bf16) : OB(PREC_BF16, ACT_SIN, __bf16);
    return act == NEB_TANH ? OB(PREC_BF16, ACT_TANH, float) : OB(PREC_BF16, ACT_SIN, float);
  }
  if (prec == NEB_F32) return act == NEB_TANH ? OB(PREC_F32, ACT_TANH, float) : OB(PREC_F32, ACT_SIN, float);
  return act == NEB_TANH ? OB(PREC_F64, ACT_TANH, double) : OB(PREC_F64, ACT_SIN, double);
#undef OB
}

int neb_jet_layer_backward(int prec, int dim, int act, int64_t n, int64_t k, int64_t width, const void* adj_in,
                           int64_t lda_in, const void* wt, int64_t ldw, const void* z_prev, int64_t ldz,
                           double omega_prev, void* adj_out, int64_t lda_out, void* stream) {
  if (int rc = check_layer(prec, dim, act)) return rc;
  if (n < 0 || k < 1 || width < 1 || lda_in < width || ldw < width || ldz < k || lda_out < k)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_backward: bad sizes");
  if (n == 0) return NEB_OK;
  if (!adj_in || !wt || !z_prev || !adj_out) return fail(NEB_ERR_ARG, "neb_jet_layer_backward: null pointer");
  if (!aligned(adj_in, lda_in, prec) || !aligned(wt, ldw, prec) || !aligned(z_prev, ldz, prec) ||
      !aligned(adj_out, lda_out, prec))
    return fail(NEB_ERR_ARG, "neb_jet_layer_backward: rows must be 16-byte aligned");
#define CALL(P, D, A) \
  layer_backward<P, D, A>(n, k, width, adj_in, lda_in, wt, ldw, z_prev, ldz, omega_prev, adj_out, lda_out, S(stream))
  DISPATCH_PDA(CALL)
#undef CALL
}

size_t neb_cross_workspace_size(int64_t m, int64_t n, int64_t rows, int symmetric) {
  int tm, tn;
  const int nt = cross_tiles(m, n, symmetric, &tm, &tn);
  const int sp = cross_splits(nt, rows);
  return (size_t)sp * nt * CR_TM * CR_TN * sizeof(double);
}

int cross_run(int mode, int nseg, const void* const* x, const void* const* y, const int64_t* ldx, const int64_t* ldy,
              const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t m, int64_t n,
              const double* roww, int symmetric, int64_t block, double scale, double beta, double* G, int64_t ldg,
              void* workspace, size_t ws_bytes, void* stream) {
  if (nseg < 1 || nseg > CR_MAXSEG) return fail(NEB_ERR_ARG, "neb_cross: 1..%d segments", CR_MAXSEG);
  if (m < 1 || n < 1 || ldg < n) return fail(NEB_ERR_SHAPE, "neb_cross: bad output shape");
  if (roww && nseg != 1) return fail(NEB_ERR_ARG, "neb_cross: row weights need a single segment");
  if (symmetric && (m != n || nseg != 1 || x[0] != y[0] || ldx[0] != ldy[0] || x_dtype != y_dtype))
    return fail(NEB_ERR_ARG, "neb_cross: symmetric mode needs X == Y");
  if (!G || !workspace) return fail(NEB_ERR_ARG, "neb_cross: null output or workspace");
  CrossArgs a;
  memset(&a, 0, sizeof(a));
  int64_t maxrows = 0;
  for (int s = 0; s < nseg; ++s) {
    if ((rows[s] > 0 && (!x[s] || !y[s])) || rows[s] < 0 || ldx[s] < m || ldy[s] < n || ldx[s] >= (1 << 26) || ldy[s] >= (1 << 26))
      return fail(NEB_ERR_SHAPE, "neb_cross: bad segment %d", s);
    a.seg[s].x = x[s]; a.seg[s].y = y[s];
    a.seg[s].ldx = ldx[s]; a.seg[s].ldy = ldy[s];
    a.seg[s].rows = rows[s]; a.seg[s].sign = sign ? sign[s] : 1.0;
    if (rows[s] > maxrows) maxrows = rows[s];
  }
  a.nseg = nseg; a.m = m; a.n = n; a.roww = roww; a.symmetric = symmetric;
  if (block > 0) {
    a.ntiles = cross_block_tiles(m, block, &a.tiles_m, &a.bdiag);
    a.tiles_n = a.tiles_m;
  } else {
    a.ntiles = cross_tiles(m, n, symmetric, &a.tiles_m, &a.tiles_n);
  }
  a.splits = cross_splits(a.ntiles, maxrows);
  const size_t need = (size_t)a.splits * a.ntiles * CR_TM * CR_TN * sizeof(double);
  if (ws_bytes < need) return fail(NEB_ERR_ARG, "neb_cross: workspace %zu < %zu bytes", ws_bytes, need);
  a.partial = reinterpret_cast<double*>(workspace);
  hipStream_t st = S(stream);
  switch (mode) {
    case NEB_CROSS_F32: return cross_dispatch_types<CR_F32>(x_dtype, y_dtype, a, scale, beta, G, ldg, st);
    case NEB_CROSS_BF16: return cross_dispatch_types<CR_BF16>(x_dtype, y_dtype, a, scale, beta, G, ldg, st);
    case NEB_CROSS_BF16X3: return cross_dispatch_types<CR_BF16X3>(x_dtype, y_dtype, a, scale, beta, G, ldg, st);
    case NEB_CROSS_F64: return cross_dispatch_types<CR_F64>(x_dtype, y_dtype, a, scale, beta, G, ldg, st);
  }
  return fail(NEB_ERR_ARG, "neb_cross: unknown mode %d", mode);
}

int neb_cross(int mode, int nseg, const void* const* x, const void* const* y, const int64_t* ldx, const int64_t* ldy,
              const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t m, int64_t n,
              const double* roww, int symmetric, double scale, double beta, double* G, int64_t ldg, void* workspace,
              size_t ws_bytes, void* stream) {
  return cross_run(mode, nseg, x, y, ldx, ldy, rows, sign, x_dtype, y_dtype, m, n, roww, symmetric, 0, scale, beta, G,
                   ldg, workspace, ws_bytes, stream);
}

size_t neb_cross_blocks_workspace_size(int64_t m, int64_t block, int64_t rows) {
  int tm, bd;
  if (m < 1 || block < 1) return 0;
  const int nt = cross_block_tiles(m, block, &tm, &bd);
  const int sp = cross_splits(nt, rows);
  return (size_t)sp * nt * CR_TM * CR_TN * sizeof(double);
}

int neb_cross_blocks(int mode, int nseg, const void* const* x, const void* const* y, const int64_t* ldx,
                     const int64_t* ldy, const int64_t* rows, const double* sign, int x_dtype, int y_dtype, int64_t m,
                     int64_t block, double scale, double beta, double* G, int64_t ldg, void* workspace,
                     size_t ws_bytes, void* stream) {
  if (block < 1 || m < 1 || m % block || (block % CR_TM != 0 && CR_TM % block != 0))
    return fail(NEB_ERR_SHAPE, "neb_cross_blocks: the block must divide m and nest with %d", CR_TM);
  return cross_run(mode, nseg, x, y, ldx, ldy, rows, sign, x_dtype, y_dtype, m, m, nullptr, 0, block, scale, beta, G,
                   ldg, workspace, ws_bytes, stream);
}

int neb_gram(int mode, const void* phi, int64_t rows, int64_t K, int64_t ld, int phi_dtype, double w,
             const double* roww, double beta, double* G, int64_t ldg, void* workspace, size_t ws_bytes,
             void* stream) {
  const void* xs[1] = {phi};
  const int64_t lds[1] = {ld};
  const int64_t rs[1] = {rows};
  const double sg[1] = {1.0};
  return neb_cross(mode, 1, xs, xs, lds, lds, rs, sg, phi_dtype, phi_dtype, K, K, roww, 1, w, beta, G, ldg,
                   workspace, ws_bytes, stream);
}

int neb_gram_split(const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld, double w, double beta,
                   double* G, int64_t ldg, void* workspace, size_t ws_bytes, void* stream) {
  if (int rc = gram_split_check("neb_gram_split", phi_hi, phi_lo, rows, K, ld, G, ldg)) return rc;
  if (!workspace) return fail(NEB_ERR_ARG, "neb_gram_split: null workspace");
  // the workspace is sized for FP64 partials (neb_cross_workspace_size)
  return gram_split_launch("neb_gram_split", phi_hi, phi_lo, rows, K, ld, w, beta, G, ldg, workspace, ws_bytes,
                           sizeof(double), stream);
}

int neb_gram_split_noalloc(const void* phi_hi, const void* phi_lo, int64_t rows, int64_t K, int64_t ld, double w,
                           double beta, double* G, int64_t ldg, void* stream) {
  if (int rc = gram_split_check("neb_gram_split_noalloc", phi_hi, phi_lo, rows, K, ld, G, ldg)) return rc;
  int tm, tn;
  const int nt = cross_tiles(K, K, 1, &tm, &tn);
  const size_t need = (size_t)cross_splits(nt, rows) * nt * CR_TM * CR_TN * sizeof(float);
  if (need > NEB_NOALLOC_SCRATCH_BYTES)
    return fail(NEB_ERR_ARG, "neb_gram_split_noalloc: K=%lld needs %zu bytes of scratch, above the %zu-byte maximum",
                (long long)K, need, (size_t)NEB_NOALLOC_SCRATCH_BYTES);
  void* ptr = nullptr;
  if (int rc = scratch_address(&ptr)) return rc;
  return gram_split_launch("neb_gram_split_noalloc", phi_hi, phi_lo, rows, K, ld, w, beta, G, ldg, ptr,
                           NEB_NOALLOC_SCRATCH_BYTES, sizeof(float), stream);
}

size_t neb_colsum_workspace_size(int64_t rows, int64_t cols) {
  return (size_t)colsum_splits(rows) * (size_t)cols * sizeof(double);
}

int neb_colsum(int dtype, const void* x, int64_t rows, int64_t cols, int64_t ld, double beta, double* out,
               void* workspace, size_t ws_bytes, void* stream) {
  if (rows < 0 || cols < 1 || ld < cols) return fail(NEB_ERR_SHAPE, "neb_colsum: bad sizes");
  if ((rows > 0 && !x) || !out || !workspace) return fail(NEB_ERR_ARG, "neb_colsum: null pointer");
  const int sp = colsum_splits(rows);
  if (ws_bytes < (size_t)sp * cols * sizeof(double)) return fail(NEB_ERR_ARG, "neb_colsum: workspace too small");
  double* part = reinterpret_cast<double*>(workspace);
  hipStream_t st = S(stream);
  switch (dtype) {
    case NEB_F32: return colsum_launch<float>(x, rows, cols, ld, beta, out, part, sp, st);
    case NEB_BF16: return colsum_launch<__bf16>(x, rows, cols, ld, beta, out, part, sp, st);
    case NEB_F64: return colsum_launch<double>(x, rows, cols, ld, beta, out, part, sp, st);
  }
  return fail(NEB_ERR_ARG, "neb_colsum: unknown dtype %d", dtype);
}

int neb_sumsq(int dtype, const void* x, int64_t n, double* out, void* workspace, size_t ws_bytes, void* stream) {
  if (n < 0 || (n > 0 && !x) || !out || !workspace) return fail(NEB_ERR_ARG, "neb_sumsq: bad arguments");
  if (ws_bytes < 256 * sizeof(double)) return fail(NEB_ERR_ARG, "neb_sumsq: workspace too small");
  double* part = reinterpret_cast<double*>(workspace);
  hipStream_t st = S(stream);
  switch (dtype) {
    case NEB_F32: hipLaunchKernelGGL(sumsq_kernel<float>, dim3(256), dim3(256), 0, st, reinterpret_cast<const float*>(x), n, part); break;
    case NEB_F64: hipLaunchKernelGGL(sumsq_kernel<double>, dim3(256), dim3(256), 0, st, reinterpret_cast<const double*>(x), n, part); break;
    case NEB_BF16: hipLaunchKernelGGL(sumsq_kernel<__bf16>, dim3(256), dim3(256), 0, st, reinterpret_cast<const __bf16*>(x), n, part); break;
    default: return fail(NEB_ERR_ARG, "neb_sumsq: unknown dtype %d", dtype);
  }
  if (int rc = hip_status("neb_sumsq")) return rc;
  hipLaunchKernelGGL(sum_splits_kernel, sum_splits_grid(1), dim3(256), 0, st, part, 256, (int64_t)1, out, 0.0);
  return hip_status("neb_sumsq reduce");
}

int neb_adam(int dtype, void* p, const void* g, void* m, void* v, int64_t n, double lr, double scale, double b1,
             double b2, double eps, int64_t t, void* stream) {
  if (n < 0 || !p || !g || !m || !v || t < 1) return fail(NEB_ERR_ARG, "neb_adam: bad arguments");
  if (n == 0) return NEB_OK;
  const double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
  const dim3 grid((unsigned)((n + 255) / 256));
  hipStream_t st = S(stream);
  if (dtype == NEB_F32)
    hipLaunchKernelGGL(adam_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<float*>(p),
                       reinterpret_cast<const float*>(g), reinterpret_cast<float*>(m), reinterpret_cast<float*>(v), n,
                       lr, scale, b1, b2, eps, c1, c2);
  else if (dtype == NEB_F64)
    hipLaunchKernelGGL(adam_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<double*>(p),
                       reinterpret_cast<const double*>(g), reinterpret_cast<double*>(m), reinterpret_cast<double*>(v),
                       n, lr, scale, b1, b2, eps, c1, c2);
  else
    return fail(NEB_ERR_ARG, "neb_adam: fp32 or fp64 master weights only");
  return hip_status("neb_adam");
}

int neb_adam_gated(int dtype, void* p, const void* g, void* m, void* v, int64_t n, double lr, double clip,
                   double b1, double b2, double eps, int64_t t, const double* norm2, const int* gate, void* stream) {
  if (n < 0 || !p || !g || !m || !v || !norm2 || t < 1 || !(clip > 0.0))
    return fail(NEB_ERR_ARG, "neb_adam_gated: bad arguments");
  if (n == 0) return NEB_OK;
  const double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
  const dim3 grid((unsigned)((n + 255) / 256));
  hipStream_t st = S(stream);
  if (dtype == NEB_F32)
    hipLaunchKernelGGL(adam_gated_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<float*>(p),
                       reinterpret_cast<const float*>(g), reinterpret_cast<float*>(m), reinterpret_cast<float*>(v), n,
                       lr, clip, b1, b2, eps, c1, c2, norm2, gate);
  else if (dtype == NEB_F64)
    hipLaunchKernelGGL(adam_gated_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<double*>(p),
                       reinterpret_cast<const double*>(g), reinterpret_cast<double*>(m), reinterpret_cast<double*>(v),
                       n, lr, clip, b1, b2, eps, c1, c2, norm2, gate);
  else
    return fail(NEB_ERR_ARG, "neb_adam_gated: fp32 or fp64 master weights only");
  return hip_status("neb_adam_gated");
}

size_t neb_spd_inverse_workspace_size(int64_t K) {
  const int64_t Kp = (K + SPD_NB - 1) / SPD_NB * SPD_NB;
  return (size_t)(2 * Kp * Kp + 4 * SPD_NB * SPD_NB) * sizeof(double);
}

// The sweeps of neb_spd_inverse; with T non-null they also write the inverse
// upper Cholesky factor (spd_step_kernel).
static int spd_sweeps(const char* name, int64_t K, const double* A, int64_t lda, double* X, int64_t ldx, double* T,
                      int64_t ldt, int* info, void* workspace, size_t ws_bytes, void* stream) {
  if (K < 1 || lda < K || ldx < K || (T && ldt < K)) return fail(NEB_ERR_SHAPE, "%s: bad sizes", name);
  if (!A || !X || !info || !workspace) return fail(NEB_ERR_ARG, "%s: null pointer", name);
  if (A == X || A == T || X == T) return fail(NEB_ERR_ARG, "%s: A, X and T must not alias", name);
  if (ws_bytes < neb_spd_inverse_workspace_size(K)) return fail(NEB_ERR_ARG, "%s: workspace too small", name);
  const int64_t nt = (K + SPD_NB - 1) / SPD_NB, Kp = nt * SPD_NB;
  if (nt > 2048) return fail(NEB_ERR_SHAPE, "%s: K too large", name);
  double* base = reinterpret_cast<double*>(workspace);
  double* buf[2] = {base, base + Kp * Kp};  // padded Kp x Kp ping-pong matrices
  double* tiles = base + 2 * Kp * Kp;       // 64 x 64 pivot inverses and T_kk^T, ping-pong
  double* Pi[2] = {tiles, tiles + SPD_NB * SPD_NB};
  double* Tk[2] = {tiles + 2 * SPD_NB * SPD_NB, tiles + 3 * SPD_NB * SPD_NB};
  hipStream_t st = S(stream);
  if (int rc = set_lds(spd_pivot0_kernel, spd_pivot0_lds_bytes())) return rc;
  if (int rc = set_lds(spd_step_kernel, spd_step_lds_bytes())) return rc;
  hipLaunchKernelGGL(spd_pivot0_kernel, dim3(1), dim3(256), spd_pivot0_lds_bytes(), st, A, lda, K, Pi[0], info,
                     T ? Tk[0] : nullptr);
  if (int rc = hip_status(name)) return rc;
  const double* src = A;
  int64_t lds = lda;
  for (int64_t k = 0; k < nt; ++k) {
    const bool last = k == nt - 1;
    double* dst = last ? X : buf[k & 1];
    const int64_t ldd = last ? ldx : Kp;
    hipLaunchKernelGGL(spd_step_kernel, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), spd_step_lds_bytes(), st,
                       src, lds, k == 0 ? 1 : 0, K, (int)k, (int)nt, Pi[k & 1], Pi[(k + 1) & 1], dst, ldd,
                       last ? 1 : 0, info, Tk[k & 1], Tk[(k + 1) & 1], T, ldt);
    if (int rc = hip_status(name)) return rc;
    src = dst;
    lds = ldd;
  }
  return NEB_OK;
}

int neb_spd_inverse(int64_t K, const double* A, int64_t lda, double* X, int64_t ldx, int* info, void* workspace,
                    size_t ws_bytes, void* stream) {
  return spd_sweeps("neb_spd_inverse", K, A, lda, X, ldx, nullptr, 0, info, workspace, ws_bytes, stream);
}

int neb_spd_inverse_chol(int64_t K, const double* A, int64_t lda, double* X, int64_t ldx, double* T, int64_t ldt,
                         int* info, void* workspace, size_t ws_bytes, void* stream) {
  if (!T) return fail(NEB_ERR_ARG, "neb_spd_inverse_chol: null pointer");
  return spd_sweeps("neb_spd_inverse_chol", K, A, lda, X, ldx, T, ldt, info, workspace, ws_bytes, stream);
}

int neb_sym_shift(int64_t K, const double* A, int64_t lda, double shift, double* B, int64_t ldb, void* stream) {
  if (K < 1 || lda < K || ldb < K || !A || !B) return fail(NEB_ERR_ARG, "neb_sym_shift: bad arguments");
  const int64_t nt = (K + SYM_T - 1) / SYM_T;
  if (nt > 65535) return fail(NEB_ERR_SHAPE, "neb_sym_shift: K too large");
  hipLaunchKernelGGL(sym_shift_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, S(stream), K, A, lda, shift,
                     B, ldb);
  return hip_status("neb_sym_shift");
}

size_t neb_spectral_workspace_size(int64_t K) {
  const int64_t nt = (K + SYM_T - 1) / SYM_T;
  return (size_t)(nt * nt * 4) * sizeof(double);
}

int neb_spectral_terms(int64_t K, const double* M, int64_t ldm, const double* Sm, int64_t lds, double* X,
                       int64_t ldx, const int* info, double lam, double ridge, double cond_limit, double* Ss,
                       int64_t ldss, double* loss, int* gate, void* workspace, size_t ws_bytes, void* stream) {
  if (K < 1 || ldm < K || lds < K || ldx < K || ldss < K) return fail(NEB_ERR_SHAPE, "neb_spectral_terms: bad sizes");
  if (!M || !Sm || !X || !Ss || !loss || !gate || !workspace)
    return fail(NEB_ERR_ARG, "neb_spectral_terms: null pointer");
  if (ws_bytes < neb_spectral_workspace_size(K)) return fail(NEB_ERR_ARG, "neb_spectral_terms: workspace too small");
  const int64_t nt = (K + SYM_T - 1) / SYM_T;
  if (nt > 65535) return fail(NEB_ERR_SHAPE, "neb_spectral_terms: K too large");
  double* part = reinterpret_cast<double*>(workspace);
  hipStream_t st = S(stream);
  hipLaunchKernelGGL(spectral_terms_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, st, K, M, ldm, Sm, lds,
                     X, ldx, Ss, ldss, ridge, part);
  if (int rc = hip_status("neb_spectral_terms")) return rc;
  hipLaunchKernelGGL(spectral_final_kernel, dim3(1), dim3(256), 0, st, part, nt * nt, info, lam, cond_limit, loss,
                     gate);
  return hip_status("neb_spectral_terms final");
}

int neb_spectral_adjoint(int64_t K, const double* C, int64_t ldc, const double* M, int64_t ldm, double lam, double* dM,
                         int64_t ldd, void* stream) {
  if (K < 1 || ldc < K || ldm < K || ldd < K || !C || !M || !dM) return fail(NEB_ERR_ARG, "neb_spectral_adjoint: bad arguments");
  if (dM == M) return fail(NEB_ERR_ARG, "neb_spectral_adjoint: dM must not alias M");
  const int64_t nt = (K + SYM_T - 1) / SYM_T;
  if (nt > 65535) return fail(NEB_ERR_SHAPE, "neb_spectral_adjoint: K too large");
  hipLaunchKernelGGL(spectral_adjoint_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, S(stream), K, C, ldc,
                     M, ldm, lam, dM, ldd);
  return hip_status("neb_spectral_adjoint");
}

int neb_split_bf16(int src_dtype, const void* src, int64_t rows, int64_t cols, int64_t lds, void* hi, void* lo,
                   int64_t ldd, void* stream) {
  if (rows < 0 || cols < 0 || lds < cols || ldd < cols) return fail(NEB_ERR_SHAPE, "neb_split_bf16: bad sizes");
  const int64_t tot = rows * cols;
  if (tot == 0) return NEB_OK;
  if (!src || !hi || !lo) return fail(NEB_ERR_ARG, "neb_split_bf16: null pointer");
  if (tot / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_split_bf16: too many elements");
  const dim3 grid((unsigned)((tot + 255) / 256));
  __bf16* h = reinterpret_cast<__bf16*>(hi);
  __bf16* l = reinterpret_cast<__bf16*>(lo);
  hipStream_t st = S(stream);
  switch (src_dtype) {
    case NEB_F32:
      hipLaunchKernelGGL(split_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<const float*>(src), rows, cols,
                         lds, h, l, ldd);
      return hip_status("neb_split_bf16");
    case NEB_F64:
      hipLaunchKernelGGL(split_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<const double*>(src), rows,
                         cols, lds, h, l, ldd);
      return hip_status("neb_split_bf16");
  }
  return fail(NEB_ERR_ARG, "neb_split_bf16: source dtype must be NEB_F32 or NEB_F64 (got %d)", src_dtype);
}

size_t neb_galerkin_workspace_size(int64_t n, int64_t K) {
  return (size_t)(2 * n + (int64_t)colsum_splits(n) * K) * sizeof(double);
}

int neb_galerkin_rhs2d(int dtype, const void* gu, const void* gv, const void* gw, int64_t n, int64_t K, int64_t ld,
                       const double* a, double w, double* rhs, void* workspace, size_t ws_bytes, void* stream) {
  if (n < 1 || K < 1 || ld < K || K > 8192) return fail(NEB_ERR_SHAPE, "neb_galerkin_rhs2d: bad sizes (K <= 8192)");
  if (!gu || !gv || !gw || !a || !rhs || !workspace) return fail(NEB_ERR_ARG, "neb_galerkin_rhs2d: null pointer");
  if (dtype != NEB_F32 && dtype != NEB_F64) return fail(NEB_ERR_ARG, "neb_galerkin_rhs2d: fields must be F32 or F64");
  const int sp = colsum_splits(n);
  if (ws_bytes < (size_t)(2 * n + (int64_t)sp * K) * sizeof(double))
    return fail(NEB_ERR_ARG, "neb_galerkin_rhs2d: workspace too small");
  double* r1 = reinterpret_cast<double*>(workspace);
  double* r2 = r1 + n;
  double* part = r2 + n;
  const int64_t chunk = (n + sp - 1) / sp;
  const dim3 gp((unsigned)((n + 3) / 4)), gc((unsigned)((K + 255) / 256), (unsigned)sp);
  const size_t lds = (size_t)K * sizeof(double);  // <= 64 KiB for K <= 8192
  hipStream_t st = S(stream);
  if (dtype == NEB_F32) {
    const float *u = reinterpret_cast<const float*>(gu), *v = reinterpret_cast<const float*>(gv),
                *o = reinterpret_cast<const float*>(gw);
    hipLaunchKernelGGL(rhs2d_points_kernel<float>, gp, dim3(256), lds, st, u, v, o, ld, n, K, a, w, r1, r2);
    if (int rc = hip_status("neb_galerkin_rhs2d")) return rc;
    hipLaunchKernelGGL(rhs2d_cols_kernel<float>, gc, dim3(256), 0, st, u, v, ld, n, K, chunk, r1, r2, part);
  } else {
    const double *u = reinterpret_cast<const double*>(gu), *v = reinterpret_cast<const double*>(gv),
                 *o = reinterpret_cast<const double*>(gw);
    hipLaunchKernelGGL(rhs2d_points_kernel<double>, gp, dim3(256), lds, st, u, v, o, ld, n, K, a, w, r1, r2);
    if (int rc = hip_status("neb_galerkin_rhs2d")) return rc;
    hipLaunchKernelGGL(rhs2d_cols_kernel<double>, gc, dim3(256), 0, st, u, v, ld, n, K, chunk, r1, r2, part);
  }
  if (int rc = hip_status("neb_galerkin_rhs2d columns")) return rc;
  hipLaunchKernelGGL(sum_splits_kernel, sum_splits_grid(K), dim3(256), 0, st, part, sp, K, rhs, 0.0);
  return hip_status("neb_galerkin_rhs2d reduce");
}

int neb_center_columns(int dtype, void* x, void* lo, int64_t rows, int64_t cols, int64_t ld, const double* mean,
                       void* stream) {
  if (rows < 0 || cols < 0 || ld < cols) return fail(NEB_ERR_SHAPE, "neb_center_columns: bad sizes");
  if (lo && dtype != NEB_BF16) return fail(NEB_ERR_ARG, "neb_center_columns: a lo plane needs dtype NEB_BF16");
  const int64_t units = rows * ((cols + 7) / 8);
  if (units == 0) return NEB_OK;
  if (!x || !mean) return fail(NEB_ERR_ARG, "neb_center_columns: null pointer");
  if (units / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_center_columns: too many elements");
  const dim3 grid((unsigned)((units + 255) / 256));
  hipStream_t st = S(stream);
  switch (dtype) {
    case NEB_F32:
      hipLaunchKernelGGL(center_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<float*>(x), nullptr, rows,
                         cols, ld, mean);
      return hip_status("neb_center_columns");
    case NEB_F64:
      hipLaunchKernelGGL(center_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<double*>(x), nullptr, rows,
                         cols, ld, mean);
      return hip_status("neb_center_columns");
    case NEB_BF16:
      hipLaunchKernelGGL(center_kernel<__bf16>, grid, dim3(256), 0, st, reinterpret_cast<__bf16*>(x),
                         reinterpret_cast<__bf16*>(lo), rows, cols, ld, mean);
      return hip_status("neb_center_columns");
  }
  return fail(NEB_ERR_ARG, "neb_center_columns: unknown dtype %d", dtype);
}

int neb_join_bf16(int dst_dtype, const void* hi, const void* lo, int64_t rows, int64_t cols, int64_t ld, void* dst,
                  int64_t ldd, void* stream) {
  if (rows < 0 || cols < 0 || ld < cols || ldd < cols) return fail(NEB_ERR_SHAPE, "neb_join_bf16: bad sizes");
  const int64_t units = rows * ((cols + 7) / 8);
  if (units == 0) return NEB_OK;
  if (!hi || !lo || !dst) return fail(NEB_ERR_ARG, "neb_join_bf16: null pointer");
  if (units / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_join_bf16: too many elements");
  const dim3 grid((unsigned)((units + 255) / 256));
  const __bf16* h = reinterpret_cast<const __bf16*>(hi);
  const __bf16* l = reinterpret_cast<const __bf16*>(lo);
  hipStream_t st = S(stream);
  switch (dst_dtype) {
    case NEB_F32:
      hipLaunchKernelGGL(join_kernel<float>, grid, dim3(256), 0, st, h, l, rows, cols, ld,
                         reinterpret_cast<float*>(dst), ldd);
      return hip_status("neb_join_bf16");
    case NEB_F64:
      hipLaunchKernelGGL(join_kernel<double>, grid, dim3(256), 0, st, h, l, rows, cols, ld,
                         reinterpret_cast<double*>(dst), ldd);
      return hip_status("neb_join_bf16");
  }
  return fail(NEB_ERR_ARG, "neb_join_bf16: destination dtype must be NEB_F32 or NEB_F64 (got %d)", dst_dtype);
}

int neb_jet_layer_forward_x3(int dim, int act, int64_t n, int64_t k, int64_t width, const void* x_hi,
                             const void* x_lo, int64_t ldx, const void* w_hi, const void* w_lo, int64_t ldw,
                             const void* bias, double omega, void* z_hi, void* z_lo, void* y_hi, void* y_lo,
                             int64_t ldy, void* stream) {
  if (int rc = check_layer(NEB_BF16, dim, act)) return rc;
  if (n < 0 || k < 1 || width < 1 || ldx < k || ldw < k || ldy < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_forward_x3: bad sizes");
  if (n == 0) return NEB_OK;
  if (!x_hi || !x_lo || !w_hi || !w_lo || !z_hi || !y_hi || !y_lo)  // z_lo may be null
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward_x3: null pointer");
  if (!aligned(x_hi, ldx, NEB_BF16) || !aligned(x_lo, ldx, NEB_BF16) || !aligned(w_hi, ldw, NEB_BF16) ||
      !aligned(w_lo, ldw, NEB_BF16) || !aligned(z_hi, ldy, NEB_BF16) || !aligned(z_lo, ldy, NEB_BF16) ||
      !aligned(y_hi, ldy, NEB_BF16) || !aligned(y_lo, ldy, NEB_BF16))
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward_x3: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
#define LFX3(D, A) \
  layer_forward_x3<D, A>(n, k, width, x_hi, x_lo, ldx, w_hi, w_lo, ldw, bias, omega, z_hi, z_lo, y_hi, y_lo, ldy, st)
  if (dim == NEB_DIM_2D_JET1) return act == NEB_TANH ? LFX3(JET1_2D, ACT_TANH) : LFX3(JET1_2D, ACT_SIN);
  if (dim == 2) return act == NEB_TANH ? LFX3(2, ACT_TANH) : LFX3(2, ACT_SIN);
  return act == NEB_TANH ? LFX3(3, ACT_TANH) : LFX3(3, ACT_SIN);
#undef LFX3
}

int neb_jet_layer_forward_x3_blocks(int dim, int act, int64_t n, int64_t width, int64_t block, const void* x_hi,
                                    const void* x_lo, int64_t ldx, const void* w_hi, const void* w_lo, int64_t ldw,
                                    const void* bias, double omega, void* z_hi, void* z_lo, void* y_hi, void* y_lo,
                                    int64_t ldy, void* stream) {
  if (int rc = check_layer(NEB_BF16, dim, act)) return rc;
  if (n < 0 || width < 1 || block < 8 || block % 8 || width % block || ldx < width || ldw < width || ldy < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_layer_forward_x3_blocks: bad sizes (block a multiple of 8 dividing the width)");
  if (n == 0) return NEB_OK;
  if (!x_hi || !x_lo || !w_hi || !w_lo || !z_hi || !y_hi || !y_lo)  // z_lo may be null
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward_x3_blocks: null pointer");
  if (!aligned(x_hi, ldx, NEB_BF16) || !aligned(x_lo, ldx, NEB_BF16) || !aligned(w_hi, ldw, NEB_BF16) ||
      !aligned(w_lo, ldw, NEB_BF16) || !aligned(z_hi, ldy, NEB_BF16) || !aligned(z_lo, ldy, NEB_BF16) ||
      !aligned(y_hi, ldy, NEB_BF16) || !aligned(y_lo, ldy, NEB_BF16))
    return fail(NEB_ERR_ARG, "neb_jet_layer_forward_x3_blocks: rows must be 16-byte aligned");
  hipStream_t st = S(stream);
#define LFX3(D, A)                                                                                                  \
  layer_forward_x3<D, A>(n, width, width, x_hi, x_lo, ldx, w_hi, w_lo, ldw, bias, omega, z_hi, z_lo, y_hi, y_lo, ldy, \
                         st, block)
  if (dim == NEB_DIM_2D_JET1) return act == NEB_TANH ? LFX3(JET1_2D, ACT_TANH) : LFX3(JET1_2D, ACT_SIN);
  if (dim == 2) return act == NEB_TANH ? LFX3(2, ACT_TANH) : LFX3(2, ACT_SIN);
  return act == NEB_TANH ? LFX3(3, ACT_TANH) : LFX3(3, ACT_SIN);
#undef LFX3
}

int neb_mode_output_fields_x3(int lift, int64_t n, int64_t h, int64_t B, const void* y_hi, const void* y_lo,
                              int64_t ldy, const void* w_hi, const void* w_lo, int64_t ldw, void* mass_hi,
                              void* mass_lo, void* stiff_hi, void* stiff_lo, int64_t ldf, void* stream) {
  if (lift != NEB_LIFT_2D_VORTICITY && lift != NEB_LIFT_3D_CURL && lift != NEB_LIFT_2D_HESSIAN && !is_stream_lift(lift))
    return fail(NEB_ERR_ARG, "neb_mode_output_fields_x3: unknown lift %d", lift);
  if (n < 0 || h < 32 || h % 32 || B < 8 || B % 8 || ldy < B * h || ldw < B * h || ldf < B)
    return fail(NEB_ERR_SHAPE, "neb_mode_output_fields_x3: bad sizes (h a multiple of 32, B a multiple of 8)");
  if (n == 0) return NEB_OK;
  if (!y_hi || !y_lo || !w_hi || !w_lo || !mass_hi || !mass_lo || !stiff_hi || !stiff_lo)
    return fail(NEB_ERR_ARG, "neb_mode_output_fields_x3: null pointer");
  if (!aligned(y_hi, ldy, NEB_BF16) || !aligned(y_lo, ldy, NEB_BF16) || !aligned(w_hi, ldw, NEB_BF16) ||
      !aligned(w_lo, ldw, NEB_BF16) || !aligned(mass_hi, ldf, NEB_BF16) || !aligned(mass_lo, ldf, NEB_BF16) ||
      !aligned(stiff_hi, ldf, NEB_BF16) || !aligned(stiff_lo, ldf, NEB_BF16))
    return fail(NEB_ERR_ARG, "neb_mode_output_fields_x3: rows must be 16-byte aligned");
  const int64_t grid = ((n + 15) / 16) * (B / 8);
  if (grid >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_mode_output_fields_x3: grid too large");
  ModeFieldsArgs a;
  a.y_hi = reinterpret_cast<const __bf16*>(y_hi); a.y_lo = reinterpret_cast<const __bf16*>(y_lo); a.ldy = ldy;
  a.w_hi = reinterpret_cast<const __bf16*>(w_hi); a.w_lo = reinterpret_cast<const __bf16*>(w_lo); a.ldw = ldw;
  a.n = n; a.h = h; a.B = B;
  a.mass_hi = reinterpret_cast<__bf16*>(mass_hi); a.mass_lo = reinterpret_cast<__bf16*>(mass_lo);
  a.stiff_hi = reinterpret_cast<__bf16*>(stiff_hi); a.stiff_lo = reinterpret_cast<__bf16*>(stiff_lo);
  a.ldf = ldf;
  hipStream_t st = S(stream);
  const dim3 g((unsigned)grid);
  if (lift == NEB_LIFT_3D_CURL) hipLaunchKernelGGL(mode_fields_kernel<3>, g, dim3(256), 0, st, a);
  else if (lift == NEB_LIFT_2D_HESSIAN) hipLaunchKernelGGL(mode_fields_kernel<4>, g, dim3(256), 0, st, a);
  else if (is_stream_lift(lift)) hipLaunchKernelGGL(mode_fields_kernel<5>, g, dim3(256), 0, st, a);
  else hipLaunchKernelGGL(mode_fields_kernel<2>, g, dim3(256), 0, st, a);
  return hip_status("neb_mode_output_fields_x3");
}

int neb_jet_output_fields_x3(int dim, int64_t n, int64_t k, int64_t K, const void* y_hi, const void* y_lo,
                             int64_t ldy, const void* w_hi, const void* w_lo, int64_t ldw, void* mass_hi,
                             void* mass_lo, void* stiff_hi, void* stiff_lo, int64_t ldf, int layout, void* stream) {
  if (int rc = check_lift(NEB_BF16, dim, NEB_TANH)) return rc;
  if (n < 0 || k < 1 || K < 1 || ldy < k || ldw < k || ldf < K || (layout != 0 && layout != 1))
    return fail(NEB_ERR_SHAPE, "neb_jet_output_fields_x3: bad sizes");
  if (n == 0) return NEB_OK;
  if (!y_hi || !y_lo || !w_hi || !w_lo || !mass_hi || !mass_lo || !stiff_hi || !stiff_lo)
    return fail(NEB_ERR_ARG, "neb_jet_output_fields_x3: null pointer");
  if (!aligned(y_hi, ldy, NEB_BF16) || !aligned(y_lo, ldy, NEB_BF16) || !aligned(w_hi, ldw, NEB_BF16) ||
      !aligned(w_lo, ldw, NEB_BF16) || !aligned(mass_hi, ldf, NEB_BF16) || !aligned(mass_lo, ldf, NEB_BF16) ||
      !aligned(stiff_hi, ldf, NEB_BF16) || !aligned(stiff_lo, ldf, NEB_BF16))
    return fail(NEB_ERR_ARG, "neb_jet_output_fields_x3: rows must be 16-byte aligned");
  return output_fields_x3(dim, n, k, K, y_hi, y_lo, ldy, w_hi, w_lo, ldw, mass_hi, mass_lo, stiff_hi, stiff_lo, ldf,
                          layout, S(stream));
}

int neb_stage_net(int master_dtype, const void* master, int nl, const int64_t* off, const int64_t* rows,
                  const int64_t* cols, void* const* hi, void* const* lo, const int64_t* ld, void* const* wt,
                  const int64_t* ldt, void* stream) {
  if (nl < 1 || nl > STAGE_NET_MAX) return fail(NEB_ERR_ARG, "neb_stage_net: 1..%d layers", STAGE_NET_MAX);
  if (!master || !off || !rows || !cols || !hi || !lo || !ld || !wt || !ldt)
    return fail(NEB_ERR_ARG, "neb_stage_net: null pointer");
  StageNetArgs a;
  memset(&a, 0, sizeof(a));
  a.nl = nl;
  for (int l = 0; l < nl; ++l) {
    if (rows[l] < 1 || cols[l] < 1 || ld[l] < cols[l] || off[l] < 0 || !hi[l] || (wt[l] && ldt[l] < rows[l]))
      return fail(NEB_ERR_SHAPE, "neb_stage_net: bad layer %d", l);
    a.off[l] = off[l]; a.rows[l] = rows[l]; a.cols[l] = cols[l];
    a.hi[l] = reinterpret_cast<__bf16*>(hi[l]);
    a.lo[l] = reinterpret_cast<__bf16*>(lo[l]);
    a.wt[l] = reinterpret_cast<__bf16*>(wt[l]);
    a.ld[l] = ld[l]; a.ldt[l] = wt[l] ? ldt[l] : 1;
    a.start_w[l + 1] = a.start_w[l] + rows[l] * ld[l];
    a.start_t[l + 1] = a.start_t[l] + (wt[l] ? cols[l] * ldt[l] : 0);
  }
  const int64_t tot = a.start_w[nl] + a.start_t[nl];
  if (tot / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_stage_net: too many elements");
  const dim3 grid((unsigned)((tot + 255) / 256));
  hipStream_t st = S(stream);
  switch (master_dtype) {
    case NEB_F32:
      hipLaunchKernelGGL(stage_net_kernel<float>, grid, dim3(256), 0, st, reinterpret_cast<const float*>(master), a);
      return hip_status("neb_stage_net");
    case NEB_F64:
      hipLaunchKernelGGL(stage_net_kernel<double>, grid, dim3(256), 0, st, reinterpret_cast<const double*>(master), a);
      return hip_status("neb_stage_net");
  }
  return fail(NEB_ERR_ARG, "neb_stage_net: master dtype must be NEB_F32 or NEB_F64 (got %d)", master_dtype);
}

int neb_stage_mode_block(int master_dtype, const void* master, int64_t K, int64_t out, int64_t in, int64_t m0,
                         int64_t B, int diag, int dst_dtype, void* hi, void* lo, int64_t ld, void* wt, int64_t ldt,
                         void* stream) {
  if (K < 1 || out < 1 || in < 1 || B < 1 || m0 < 0 || m0 + B > K || diag < 0 || diag > 2)
    return fail(NEB_ERR_SHAPE, "neb_stage_mode_block: bad sizes");
  const int64_t rows = B * out, cols = diag ? B * in : in;
  if (ld < cols || (wt && ldt < rows)) return fail(NEB_ERR_SHAPE, "neb_stage_mode_block: bad strides");
  if (!master || !hi) return fail(NEB_ERR_ARG, "neb_stage_mode_block: null pointer");
  if (lo && dst_dtype != NEB_BF16) return fail(NEB_ERR_ARG, "neb_stage_mode_block: lo planes are bf16 only");
  const int64_t tot = diag == 2 ? 2 * B * out * in : rows * ld + (wt ? cols * ldt : 0);
  if (tot / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_stage_mode_block: too many elements");
  const dim3 grid((unsigned)((tot + 255) / 256));
  hipStream_t st = S(stream);
#define SMB(TS, TD)                                                                                              \
  hipLaunchKernelGGL((stage_modes_kernel<TS, TD>), grid, dim3(256), 0, st, reinterpret_cast<const TS*>(master), out, \
                     in, m0, B, diag, reinterpret_cast<TD*>(hi), reinterpret_cast<TD*>(lo), ld,                    \
                     reinterpret_cast<TD*>(wt), ldt);                                                             \
  return hip_status("neb_stage_mode_block");
  if (master_dtype == NEB_F32) {
    if (dst_dtype == NEB_BF16) { SMB(float, __bf16) }
    if (dst_dtype == NEB_F32) { SMB(float, float) }
    if (dst_dtype == NEB_F64) { SMB(float, double) }
  } else if (master_dtype == NEB_F64) {
    if (dst_dtype == NEB_BF16) { SMB(double, __bf16) }
    if (dst_dtype == NEB_F32) { SMB(double, float) }
    if (dst_dtype == NEB_F64) { SMB(double, double) }
  }
#undef SMB
  return fail(NEB_ERR_ARG, "neb_stage_mode_block: unsupported dtypes %d -> %d", master_dtype, dst_dtype);
}

int neb_stage(int src_dtype, int dst_dtype, const void* src, int64_t rows, int64_t cols, int64_t lds, void* dst,
              int64_t ldd, int64_t drows, int64_t dcols, int transpose, double scale, void* stream) {
  if (!src || !dst || rows < 0 || cols < 0 || drows < 0 || dcols < 0 || ldd < dcols)
    return fail(NEB_ERR_ARG, "neb_stage: bad arguments");
  hipStream_t st = S(stream);
  switch (src_dtype) {
    case NEB_F32: return stage_dst<float>(dst_dtype, src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
    case NEB_F64: return stage_dst<double>(dst_dtype, src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
    case NEB_BF16: return stage_dst<__bf16>(dst_dtype, src, rows, cols, lds, dst, ldd, drows, dcols, transpose, scale, st);
  }
  return fail(NEB_ERR_ARG, "neb_stage: unknown src dtype %d", src_dtype);
}

int neb_stack_channels(const void* src_hi, const void* src_lo, int64_t n, int64_t ld, int64_t width, int nch,
                       const int* ch, void* dst_hi, void* dst_lo, int64_t ldd, void* stream) {
  if (n < 0 || width < 8 || width % 8 || ld < width || nch < 1 || nch > STACK_MAXCH || ldd < (int64_t)nch * width)
    return fail(NEB_ERR_SHAPE, "neb_stack_channels: bad sizes (width a multiple of 8, 1 <= nch <= %d)", STACK_MAXCH);
  if (n == 0) return NEB_OK;
  if (!src_hi || !dst_hi || !ch || (!src_lo) != (!dst_lo)) return fail(NEB_ERR_ARG, "neb_stack_channels: null pointer");
  if (!aligned(src_hi, ld, NEB_BF16) || !aligned(dst_hi, ldd, NEB_BF16) || (src_lo && !aligned(src_lo, ld, NEB_BF16)) ||
      (dst_lo && !aligned(dst_lo, ldd, NEB_BF16)))
    return fail(NEB_ERR_ARG, "neb_stack_channels: rows must be 16-byte aligned");
  ChanMap map;
  memset(&map, 0, sizeof(map));
  map.n = nch;
  for (int i = 0; i < nch; ++i) {
    if (ch[i] < 0 || ch[i] >= RG_MAXCH) return fail(NEB_ERR_ARG, "neb_stack_channels: channel %d out of range", ch[i]);
    map.ch[i] = ch[i];
  }
  const int64_t units = n * nch * (width / 8);
  if (units == 0) return NEB_OK;
  if (units / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_stack_channels: too many elements");
  hipLaunchKernelGGL(stack_channels_kernel, dim3((unsigned)((units + 255) / 256)), dim3(256), 0, S(stream),
                     reinterpret_cast<const __bf16*>(src_hi), reinterpret_cast<const __bf16*>(src_lo), n, ld, width,
                     map, reinterpret_cast<__bf16*>(dst_hi), reinterpret_cast<__bf16*>(dst_lo), ldd);
  return hip_status("neb_stack_channels");
}

int neb_jet_act_reverse(int prec, int dim, int act, int64_t n, int64_t width, const float* adj, int64_t lda, int nch,
                        const int* ch, const void* z, int64_t ldz, double omega, void* out, int64_t ldo,
                        void* stream) {
  if (int rc = check_common(prec, dim, act)) return rc;
  if (prec == NEB_F64) return fail(NEB_ERR_ARG, "neb_jet_act_reverse: tapes are BF16 or F32");
  const int C = dim == 2 ? 6 : 10;
  if (n < 0 || width < 1 || width % 4 || nch < 1 || nch > STACK_MAXCH || lda < (int64_t)nch * width || ldz < width ||
      ldo < width)
    return fail(NEB_ERR_SHAPE, "neb_jet_act_reverse: bad sizes (width a multiple of 4, 1 <= nch <= %d)", STACK_MAXCH);
  if (n == 0) return NEB_OK;
  if (!adj || !ch || !z || !out) return fail(NEB_ERR_ARG, "neb_jet_act_reverse: null pointer");
  if (!aligned(adj, lda, NEB_F32) || !aligned(z, ldz, prec) || !aligned(out, ldo, prec))
    return fail(NEB_ERR_ARG, "neb_jet_act_reverse: rows must be 16-byte aligned");
  ChanMap map;
  memset(&map, 0, sizeof(map));
  map.n = nch;
  for (int i = 0; i < nch; ++i) {
    if (ch[i] < 0 || ch[i] >= C) return fail(NEB_ERR_ARG, "neb_jet_act_reverse: channel %d out of range", ch[i]);
    map.ch[i] = ch[i];
  }
  const int64_t units = n * ((width + 7) / 8);
  if (units == 0) return NEB_OK;
  if (units / 256 >= 0x7fffffff) return fail(NEB_ERR_SHAPE, "neb_jet_act_reverse: too many elements");
  const dim3 grid((unsigned)((units + 255) / 256));
  hipStream_t st = S(stream);
#define ARS(D, A, T)                                                                                              \
  hipLaunchKernelGGL((act_reverse_stacked_kernel<D, A, T>), grid, dim3(256), 0, st, adj, lda, map,               \
                     reinterpret_cast<const T*>(z), ldz, n, width, omega, reinterpret_cast<T*>(out), ldo);      \
  return hip_status("neb_jet_act_reverse");
  if (prec == NEB_BF16) {
    if (dim == 2) { if (act == NEB_TANH) { ARS(2, ACT_TANH, __bf16) } else { ARS(2, ACT_SIN, __bf16) } }
    if (act == NEB_TANH) { ARS(3, ACT_TANH, __bf16) } else { ARS(3, ACT_SIN, __bf16) }
  }
  if (dim == 2) { if (act == NEB_TANH) { ARS(2, ACT_TANH, float) } else { ARS(2, ACT_SIN, float) } }
  if (act == NEB_TANH) { ARS(3, ACT_TANH, float) } else { ARS(3, ACT_SIN, float) }
#undef ARS
}

int neb_gemm_split(int64_t n, int64_t k, int64_t nout, const void* x_hi, const void* x_lo, int64_t ldx,
                   const void* w_hi, const void* w_lo, int64_t ldw, float* out, int64_t ldo, void* stream) {
  if (n < 0 || k < 1 || nout < 1 || ldx < k || ldw < k || ldo < nout)
    return fail(NEB_ERR_SHAPE, "neb_gemm_split: bad sizes");
  if (n == 0) return NEB_OK;
  if (!x_hi || !x_lo || !w_hi || !w_lo || !out) return fail(NEB_ERR_ARG, "neb_gemm_split: null pointer");
  if (!aligned(x_hi, ldx, NEB_BF16) || !aligned(x_lo, ldx, NEB_BF16) || !aligned(w_hi, ldw, NEB_BF16) ||
      !aligned(w_lo, ldw, NEB_BF16) || !aligned(out, ldo, NEB_F32))
    return fail(NEB_ERR_ARG, "neb_gemm_split: rows must be 16-byte aligned");
  RowArgs a = row_args();
  a.x[0] = x_hi; a.ldx[0] = ldx;
  a.x[1] = x_lo; a.ldx[1] = ldx;
  a.w[0] = w_hi; a.ldw[0] = ldw;
  a.w[1] = w_lo; a.ldw[1] = ldw;
  a.npts = n; a.nout = nout; a.k = k;
  a.out0 = out; a.ldo = ldo;
  return launch_rowgemm<PREC_BF16, __bf16, PlanGemmSplit>(a, S(stream), "neb_gemm_split");
}

}  // extern "C"
