// Common types for the gfx950 (CDNA4) neural-eigenfluid kernels.
//
// Precision policies select the MFMA instruction used by every contraction:
//   PREC_BF16 : bf16 operands, v_mfma_f32_16x16x32_bf16, fp32 accumulate
//   PREC_F32  : fp32 operands, v_mfma_f32_16x16x4f32
//   PREC_F64  : fp64 operands, v_mfma_f64_16x16x4f64 (finite-difference checks)
//
// All MFMA tiles are 16x16 with a 64-wide wavefront.  Operand fragments are
// read from LDS tiles stored [row][k] (k contiguous): lane l reads row l%16 and
// the KE consecutive k values starting at (l/16)*KE.  A and B fragments use the
// same lane->k map, so the contraction is correct for any hardware k order.
// The 16x16 result is held as lane l, element r -> column l%16 and row
// drow(l, r): 4*(l/16)+r for the fp32-accumulating instructions, (l/16)+4*r
// for v_mfma_f64_16x16x4f64 (measured on gfx950).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

enum { PREC_F32 = 0, PREC_BF16 = 1, PREC_F64 = 2 };
enum { ACT_TANH = 0, ACT_SIN = 1 };

template <int P> struct Prec;

template <> struct Prec<PREC_BF16> {
  using store_t = __bf16;  // activations / tape in HBM and MFMA operands in LDS
  using math_t = float;    // epilogue arithmetic
  using frag_t = bf16x8;
  using acc_t = f32x4;
  using accel_t = float;
  static constexpr int KI = 32;  // k per MFMA
  static constexpr int KE = 8;   // k elements per lane per fragment
  static constexpr int RK = 32;  // k per LDS stage
  static constexpr int PAD = 0;  // unpadded rows, chunk-swizzled (lds_k)
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static inline int rowg(int rho) { return rho >> 2; }  // inverse of drow
  __device__ static inline int rowr(int rho) { return rho & 3; }
};

template <> struct Prec<PREC_F32> {
  using store_t = float;
  using math_t = float;
  using frag_t = float;
  using acc_t = f32x4;
  using accel_t = float;
  static constexpr int KI = 4;
  static constexpr int KE = 1;
  static constexpr int RK = 32;
  static constexpr int PAD = 4;
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return 4 * (lane >> 4) + r; }
  __device__ static inline int rowg(int rho) { return rho >> 2; }  // inverse of drow
  __device__ static inline int rowr(int rho) { return rho & 3; }
};

template <> struct Prec<PREC_F64> {
  using store_t = double;
  using math_t = double;
  using frag_t = double;
  using acc_t = f64x4;
  using accel_t = double;
  static constexpr int KI = 4;
  static constexpr int KE = 1;
  static constexpr int RK = 16;
  static constexpr int PAD = 2;
  __device__ static inline acc_t mfma(frag_t a, frag_t b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int drow(int lane, int r) { return (lane >> 4) + 4 * r; }
  __device__ static inline int rowg(int rho) { return rho & 3; }
  __device__ static inline int rowr(int rho) { return rho >> 2; }
};

// ---------------------------------------------------------------------------
// scalar conversions
// ---------------------------------------------------------------------------
template <typename T> __device__ inline T cvt(float v);
template <typename T> __device__ inline T cvt(double v);
template <> __device__ inline float cvt<float>(float v) { return v; }
template <> __device__ inline float cvt<float>(double v) { return (float)v; }
template <> __device__ inline double cvt<double>(float v) { return (double)v; }
template <> __device__ inline double cvt<double>(double v) { return v; }
template <> __device__ inline __bf16 cvt<__bf16>(float v) { return (__bf16)v; }
template <> __device__ inline __bf16 cvt<__bf16>(double v) { return (__bf16)(float)v; }

template <typename T> __device__ inline double todbl_any(T v) { return (double)v; }
template <> __device__ inline double todbl_any<__bf16>(__bf16 v) { return (double)(float)v; }

template <typename M, typename T> __device__ inline M tomath(T v) { return (M)v; }
template <> __device__ inline float tomath<float, __bf16>(__bf16 v) { return (float)v; }
template <> __device__ inline double tomath<double, __bf16>(__bf16 v) { return (double)(float)v; }

// ---------------------------------------------------------------------------
// activations: value and first three derivatives of f(u)
// ---------------------------------------------------------------------------
template <int ACT, typename M> struct Act;

template <typename M> struct Act<ACT_TANH, M> {
  __device__ static inline void eval(M u, M& f0, M& f1, M& f2) {
    M t = tanh(u);
    M d = (M)1 - t * t;
    f0 = t; f1 = d; f2 = (M)-2 * t * d;
  }
  __device__ static inline void eval3(M u, M& f1, M& f2, M& f3) {
    M t = tanh(u);
    M d = (M)1 - t * t;
    f1 = d; f2 = (M)-2 * t * d; f3 = (M)-2 * d * ((M)1 - (M)3 * t * t);
  }
};

template <typename M> struct Act<ACT_SIN, M> {
  __device__ static inline void eval(M u, M& f0, M& f1, M& f2) {
    M s = sin(u), c = cos(u);
    f0 = s; f1 = c; f2 = -s;
  }
  __device__ static inline void eval3(M u, M& f1, M& f2, M& f3) {
    M s = sin(u), c = cos(u);
    f1 = c; f2 = -s; f3 = -c;
  }
};

// Jet channel layouts, keyed by the kernels' dimension code D: C packed
// channels (value, the S spatial gradients, then P Hessian pairs in the
// reference order: 2D (0,0),(0,1),(1,1); 3D (0,0),(0,1),(0,2),(1,1),(1,2),
// (2,2)).  D = JET1_2D: 2D first-order jets (value and gradient only, the
// Jet1 stream of HB2P:1716-1800).
constexpr int JET1_2D = 12;
template <int D> struct Jet2Layout;
template <> struct Jet2Layout<2> {
  static constexpr int C = 6, P = 3, S = 2;
  __device__ static constexpr int pi(int p) { return p == 2 ? 1 : 0; }
  __device__ static constexpr int pj(int p) { return p == 0 ? 0 : 1; }
};
template <> struct Jet2Layout<3> {
  static constexpr int C = 10, P = 6, S = 3;
  __device__ static constexpr int pi(int p) { return p < 3 ? 0 : (p < 5 ? 1 : 2); }
  __device__ static constexpr int pj(int p) { return p < 3 ? p : (p < 5 ? p - 2 : 2); }
};
template <> struct Jet2Layout<JET1_2D> {
  static constexpr int C = 3, P = 0, S = 2;
  __device__ static constexpr int pi(int) { return 0; }
  __device__ static constexpr int pj(int) { return 0; }
};

// Second-order jet product y = e x (the envelope product rule, VJ2:47-50):
//   y = e x,  grad y = e grad x + x grad e,
//   hess_ij y = e hess_ij x + x hess_ij e + e_i x_j + e_j x_i.
template <int D, typename M>
__device__ inline void jet_product(const M* e, const M* x, M* y) {
  using L = Jet2Layout<D>;
  y[0] = e[0] * x[0];
#pragma unroll
  for (int i = 0; i < D; ++i) y[1 + i] = e[0] * x[1 + i] + e[1 + i] * x[0];
#pragma unroll
  for (int q = 0; q < L::P; ++q) {
    const int i = L::pi(q), j = L::pj(q);
    y[1 + D + q] = e[0] * x[1 + D + q] + e[1 + D + q] * x[0] + e[1 + i] * x[1 + j] + e[1 + j] * x[1 + i];
  }
}

// Its transpose in x (VJ2:51-55): r = d<a, e x>/dx for the cotangent a of y.
template <int D, typename M>
__device__ inline void jet_product_reverse(const M* e, const M* a, M* r) {
  using L = Jet2Layout<D>;
  r[0] = e[0] * a[0];
#pragma unroll
  for (int i = 0; i < D; ++i) {
    r[0] += e[1 + i] * a[1 + i];
    r[1 + i] = e[0] * a[1 + i];
  }
#pragma unroll
  for (int q = 0; q < L::P; ++q) {
    const int i = L::pi(q), j = L::pj(q);
    r[0] += e[1 + D + q] * a[1 + D + q];
    r[1 + i] += e[1 + j] * a[1 + D + q];
    r[1 + j] += e[1 + i] * a[1 + D + q];
    r[1 + D + q] = e[0] * a[1 + D + q];
  }
}

// Forward second-order chain rule for y = f(u) on one (point, feature) pair.
// in/out: [C] packed channels (value, grads, hessian pairs).
template <int D, int ACT, typename M>
__device__ inline void act_jet_forward(const M* u, M* y) {
  using L = Jet2Layout<D>;
  constexpr int S = L::S;
  M f0, f1, f2;
  Act<ACT, M>::eval(u[0], f0, f1, f2);
  y[0] = f0;
#pragma unroll
  for (int a = 0; a < S; ++a) y[1 + a] = f1 * u[1 + a];
#pragma unroll
  for (int p = 0; p < L::P; ++p)
    y[1 + S + p] = f1 * u[1 + S + p] + f2 * u[1 + L::pi(p)] * u[1 + L::pj(p)];
}

// Transpose of act_jet_forward w.r.t. the pre-activation jet u given the
// output adjoint a; result in b.  Off-diagonal pairs contribute to both grads.
template <int D, int ACT, typename M>
__device__ inline void act_jet_reverse(const M* u, const M* a, M* b) {
  using L = Jet2Layout<D>;
  constexpr int S = L::S;
  M f1, f2, f3;
  Act<ACT, M>::eval3(u[0], f1, f2, f3);
  M bv = a[0] * f1;
#pragma unroll
  for (int c = 0; c < S; ++c) {
    bv += f2 * a[1 + c] * u[1 + c];
    b[1 + c] = f1 * a[1 + c];
  }
#pragma unroll
  for (int p = 0; p < L::P; ++p) {
    const int i = L::pi(p), j = L::pj(p);
    const M ah = a[1 + S + p];
    bv += ah * (f2 * u[1 + S + p] + f3 * u[1 + i] * u[1 + j]);
    b[1 + i] += f2 * ah * u[1 + j];
    b[1 + j] += f2 * ah * u[1 + i];
    b[1 + S + p] = f1 * ah;
  }
  b[0] = bv;
}

// LDS position of k-offset k (a multiple of 8 for bf16) inside tile row
// `row`.  bf16 tiles hold RK = 32 k-values per row, four 16-byte chunks,
// unpadded; chunk c of row r sits at slot c ^ ((r >> SH) & 3).  On gfx950 a
// ds_read_b128 serves 16 lanes per cycle as the lane blocks {8a..8a+7,
// 8b..8b+7} with (a, b) in {(0,3), (1,2), (4,7), (5,6)} over 64 banks, and a
// ds_write_b128 serves 8 consecutive lanes over 32 banks (measured with
// SQ_LDS_BANK_CONFLICT).  The row-GEMM (SH = 2: its loader writes 2 rows x
// 4 chunks per 8 lanes, its weight fragments read rows 0-3, 8-11, 16-19,
// 24-27) and the cross GEMM (SH = 1: its loader writes 8 consecutive rows x
// one chunk per 8 lanes) are then bank-conflict free for both the MFMA
// fragment reads and the loader writes, where padded rows made one or the
// other 2-way.  fp32 / fp64 tiles keep padded rows (k unchanged).
template <typename S, int SH>
__device__ inline int lds_k(int row, int k) {
  if constexpr (sizeof(S) == 2) return ((((k >> 3) ^ (row >> SH)) & 3) << 3) | (k & 7);
  return k;
}

// Makes x opaque to the optimiser (kept in a VGPR, value unknown to it).
// The asm exists in the device pass only: the host pass of a kernel body
// does not know the AMDGPU register constraint.
template <typename T>
__device__ inline void opaque_vgpr(T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(x));
#endif
}

// s_waitcnt vmcnt(n) with expcnt / lgkmcnt left at their maxima (vmcnt is
// split over bits 3:0 and 15:14 of the immediate)
constexpr int vmcnt_imm(int n) { return 0x0F70 | (n & 15) | ((n >> 4) << 14); }
// this wave waits for all but the last `later` stages of copies (LPW each)
template <int LPW, int NS>
__device__ inline void wait_copies(int later) {
  static_assert(LPW * (NS - 2) <= 63, "vmcnt range");
  if (later >= NS - 2) __builtin_amdgcn_s_waitcnt(vmcnt_imm(LPW * (NS - 2)));
  else if (NS > 3 && later == 1) __builtin_amdgcn_s_waitcnt(vmcnt_imm(LPW));
  else if (later == 0) __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));
  else __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // deeper rings: conservative
}

// Workgroup barrier for the asynchronous-copy rings.  __syncthreads() is a
// release/acquire fence plus s_barrier, and the compiler implements the
// fence as s_waitcnt vmcnt(0): every copy still in flight (global -> LDS
// copies count against vmcnt) would drain at each barrier.  The rings order
// their copies with explicit vmcnt waits instead, so this barrier only
// completes the wave's own LDS accesses (lgkmcnt(0)); the empty asm
// statements keep the compiler from moving memory accesses across it.
__device__ inline void ring_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0), vmcnt / expcnt untouched
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// A list of packed jet channels (the stacked-channel kernels: stack_channels,
// act_reverse_stacked), passed by value.
constexpr int STACK_MAXCH = 16;
struct ChanMap {
  int n;
  int ch[STACK_MAXCH];
};
