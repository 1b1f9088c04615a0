// Output layer + lift of a block of per-mode networks on split operands
// (the reference's K independent SIRENs, HB2P:1590-1880 / NCC:111-185; the
// mode's output  A_m = Y_m W_m^T  and its curl / velocity / streamfunction
// lift, NCC:97-99, VJ2:93-101, ELC:35-38, HB2P:3022, HB2P:3461-3474).
//
// In a block of B modes the output weights are block-diagonal: mode m's
// ncomp (1 in 2D, 3 in 3D) output rows read only its own h hidden
// features.  A feature tile of the trunk's output kernel spans several modes
// and would contract all B*h inputs (B times the useful work), so this
// kernel contracts each mode on its own: one v_mfma_f32_16x16x32_bf16 per
// (channel, 32-feature k-step, split product) with the mode's ncomp weight
// rows as the MFMA's A rows (rows ncomp..15 zero) and 16 points as its
// columns -- 3/16 of the MFMA is useful, but the contraction is h long, not
// B*h, and the work is a small fraction of the hidden layers'.
//
// Lanes: a workgroup owns 16 points and 8 consecutive modes; each of its 4
// waves contracts 2 of the modes (lane l holds output rows 4(l/16)+r of
// point l%16: the ncomp outputs sit in lanes 0..15), applies the lift in
// those lanes and writes, per field row, the 2 modes' columns as one 4-byte
// hi and one 4-byte lo store (the split field planes: hi = bf16(x),
// lo = bf16(x - hi)); the 4 waves fill each row's 16-byte column group.
// Four waves per 16 points keep enough loads in flight: the contraction is
// latency-bound, not MFMA-bound.
#pragma once
#include <type_traits>
#include "jet_rowgemm.hpp"

struct ModeFieldsArgs {
  const __bf16* y_hi;     // block's last hidden jets, packed [C][n][ldy]
  const __bf16* y_lo;
  int64_t ldy;
  const __bf16* w_hi;     // block-diagonal output weights [B*ncomp][ldw]: row m*ncomp+c,
  const __bf16* w_lo;     // mode m's h weights at columns m*h .. m*h+h-1
  int64_t ldw;
  int64_t n, h, B;
  __bf16* mass_hi;        // field planes, point-major rows, columns 0..B-1
  __bf16* mass_lo;
  __bf16* stiff_hi;
  __bf16* stiff_lo;
  int64_t ldf;
};

// LIFT: NEB_LIFT_2D_VORTICITY (2), NEB_LIFT_3D_CURL (3), NEB_LIFT_2D_HESSIAN (4),
// NEB_LIFT_2D_STREAM (5, also for first-order tapes).  CU: the jet channels
// the lift reads (the first CU of the tape), all of them unless set.
template <int LIFT> struct ModeLift;
template <class L, class = void> struct ModeLiftChannels : std::integral_constant<int, Jet2Layout<L::D>::C> {};
template <class L>
struct ModeLiftChannels<L, std::void_t<decltype(L::CU)>> : std::integral_constant<int, L::CU> {};
template <> struct ModeLift<2> {
  static constexpr int D = 2, NC = 1, CM = 2, CS = 1;
  // a[c][comp]: raw output jet channel c of component comp
  __device__ static inline void apply(const float (*a)[3], float* m, float* s) {
    m[0] = a[2][0];                       // u = psi_y
    m[1] = -a[1][0];                      // v = -psi_x
    s[0] = -(a[3][0] + a[5][0]);          // omega = -(psi_xx + psi_yy)
  }
};
template <> struct ModeLift<4> {
  static constexpr int D = 2, NC = 1, CM = 2, CS = 3;
  __device__ static inline void apply(const float (*a)[3], float* m, float* s) {
    m[0] = a[2][0];
    m[1] = -a[1][0];
    s[0] = a[3][0];                       // psi_xx, sqrt2 psi_xy, psi_yy
    s[1] = SQRT2F * a[4][0];
    s[2] = a[5][0];
  }
};
template <> struct ModeLift<5> {
  static constexpr int D = 2, NC = 1, CM = 1, CS = 2, CU = 3;
  __device__ static inline void apply(const float (*a)[3], float* m, float* s) {
    m[0] = a[0][0];                       // psi (no bias: the peel's centering removes it)
    s[0] = a[1][0];                       // psi_x, psi_y
    s[1] = a[2][0];
  }
};
template <> struct ModeLift<3> {
  static constexpr int D = 3, NC = 3, CM = 3, CS = 9;
  __device__ static inline int hch(int i, int j) {  // Hessian channel of the pair (i, j)
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    return 4 + (lo == 0 ? hi : (lo == 1 ? 2 + hi : 5));
  }
  __device__ static inline void apply(const float (*a)[3], float* m, float* s) {
    // u_a = d_{a+1} A_{a+2} - d_{a+2} A_{a+1};  d_j u_a from the Hessian channels
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int q1 = (q + 1) % 3, q2 = (q + 2) % 3;
      m[q] = a[1 + q1][q2] - a[1 + q2][q1];
#pragma unroll
      for (int j = 0; j < 3; ++j) s[3 * q + j] = a[hch(j, q1)][q2] - a[hch(j, q2)][q1];
    }
  }
};

template <int LIFT>
__global__ __launch_bounds__(256) void mode_fields_kernel(ModeFieldsArgs a) {
  using L = ModeLift<LIFT>;
  constexpr int C = ModeLiftChannels<L>::value, NC = L::NC, CM = L::CM, CS = L::CS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int fr = lane & 15, g = lane >> 4;
  const int64_t mgroups = a.B / 8;
  const int64_t p0 = ((int64_t)blockIdx.x / mgroups) * 16;  // 16-point tile
  const int64_t m2 = ((int64_t)blockIdx.x % mgroups) * 8 + 2 * wave;  // the wave's first mode
  const int64_t p = p0 + fr;
  const int64_t pl = p < a.n ? p : a.n - 1;                  // rows past n re-read the last (not stored)
  const int nks = (int)(a.h / 32);
  const int64_t cs = a.n * a.ldy;
  const bool wrow = fr < NC;

  float fm[2][CM], fs[2][CS];
#pragma unroll
  for (int mi = 0; mi < 2; ++mi) {
    const int64_t m = m2 + mi;
    f32x4 acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = f32x4{0.f, 0.f, 0.f, 0.f};
    const __bf16* wh = a.w_hi + (m * NC + (wrow ? fr : 0)) * a.ldw + m * a.h + 8 * g;
    const __bf16* wl = a.w_lo + (m * NC + (wrow ? fr : 0)) * a.ldw + m * a.h + 8 * g;
    const __bf16* yh = a.y_hi + pl * a.ldy + m * a.h + 8 * g;
    const __bf16* yl = a.y_lo + pl * a.ldy + m * a.h + 8 * g;
    for (int ks = 0; ks < nks; ++ks) {
      bf16x8 w0 = *reinterpret_cast<const bf16x8*>(wh + 32 * ks);
      bf16x8 w1 = *reinterpret_cast<const bf16x8*>(wl + 32 * ks);
      if (!wrow) {
        w0 = bf16x8{};
        w1 = bf16x8{};
      }
      bf16x8 xh[C], xl[C];
#pragma unroll
      for (int c = 0; c < C; ++c) {
        xh[c] = *reinterpret_cast<const bf16x8*>(yh + c * cs + 32 * ks);
        xl[c] = *reinterpret_cast<const bf16x8*>(yl + c * cs + 32 * ks);
      }
#pragma unroll
      for (int c = 0; c < C; ++c) {
        // small products first: W_lo x_hi + W_hi x_lo + W_hi x_hi
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w1, xh[c], acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, xl[c], acc[c], 0, 0, 0);
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w0, xh[c], acc[c], 0, 0, 0);
      }
    }
    // lanes 0..15 hold output rows 0..3 of point fr: the ncomp outputs
    float raw[C][3];
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
      for (int r = 0; r < 3; ++r) raw[c][r] = acc[c][r];
    L::apply(raw, fm[mi], fs[mi]);
  }
  if (g != 0 || p >= a.n) return;
#pragma unroll
  for (int q = 0; q < CM; ++q) {
    const float v[2] = {fm[0][q], fm[1][q]};
    st_field<true, 2, float>(a.mass_hi, a.mass_lo, (p * CM + q) * a.ldf + m2, v, 2);
  }
#pragma unroll
  for (int q = 0; q < CS; ++q) {
    const float v[2] = {fs[0][q], fs[1][q]};
    st_field<true, 2, float>(a.stiff_hi, a.stiff_lo, (p * CS + q) * a.ldf + m2, v, 2);
  }
}
