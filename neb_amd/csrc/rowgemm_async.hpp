// Persistent row GEMM for bf16 operands fed by asynchronous global -> LDS
// copies: the same plans, accumulator sets, fragment permutation and
// epilogues as rowgemm_kernel (jet_rowgemm.hpp), with a different data path.
//
//   * global_load_lds_dwordx4 copies 16 bytes per lane straight into LDS
//     (no staging VGPRs); one copy instruction fills 16 tile rows x 32 k
//     (1 KB).  The row-major sources are k-contiguous, which is the MFMA
//     operand order, so the tiles keep the chunk-swizzled layout of
//     lds_k<S, 2> and the fragments are plain ds_read_b128.
//   * A ring of NS LDS slots (one LDS object each, so the compiler's wait
//     insertion tracks the copies per slot) keeps NS-1 k-steps in flight.
//   * Each workgroup walks a contiguous range of (point block, feature block)
//     tiles; the copy cursor runs ahead across tile boundaries, so the next
//     tile's first k-steps land while the current tile's epilogue (the jet
//     chain and stores) runs.
//   * A wait for a load waits for every earlier copy and store of the wave
//     (the counter is in issue order), so epilogues load nothing after the
//     copies of the step: a bias (value channel) is held in registers from
//     the start, and the tape of a reverse epilogue is loaded for the whole
//     tile before the last k-step's refill, whose copies then stay in flight
//     through the epilogue.
// Rows past the point / feature count re-read the last row (their results
// are never stored); k past the reduction length is zeroed in LDS on the
// last k-step of a tile.
#pragma once
#include <type_traits>

#include "jet_rowgemm.hpp"

// Plan members the async kernel looks for: BIAS_CH0 (the epilogue adds an
// fp32 bias to the value channel) and TAPE_CH (the epilogue reads that many
// tape channels at its own (p, f) group and has an epilogue_tape that takes
// them loaded).
template <class Plan, class = void> struct PlanBiasCh0 : std::false_type {};
template <class Plan>
struct PlanBiasCh0<Plan, std::void_t<decltype(Plan::BIAS_CH0)>> : std::bool_constant<Plan::BIAS_CH0> {};
template <class Plan, class = void> struct PlanTapeCh : std::integral_constant<int, 0> {};
template <class Plan>
struct PlanTapeCh<Plan, std::void_t<decltype(Plan::TAPE_CH)>> : std::integral_constant<int, Plan::TAPE_CH> {};

// Tile shape of the async launch: the plan's own, except for the hidden
// layers, whose 256-feature width fits one workgroup's tile (2D: 8 waves, 4
// across 64 features each x 2 across 16 points each, a 256 x 32 tile; 3D:
// 32 features per wave, 128 x 32, half the width).  A point block split over
// feature tiles is read once per tile, and the workgroups holding the other
// tiles of a block drift further apart than an XCD's L2 holds (the hidden
// layers read their input jets twice from HBM with two 128-feature tiles);
// one tile across the width reads it once.  The accumulators per wave stay
// what they were (the same 64 or 32 features x 16 points of every channel).
// Eight waves, one workgroup per CU with four 28 KB slots, keep 1.5x the
// points in flight of two 4-wave workgroups with three 22.5 KB slots each,
// and re-read the weight tile from L2 once per 32 points instead of 16
// (2D forward 209 -> 198 us, reverse 206 -> 196 us).
template <class Plan>
struct RowAsyncShape {
  static constexpr int FB = Plan::FB, PB = Plan::PB, WPN = Plan::WPN, WPP = Plan::WPP;
};
template <int D> struct WideHiddenShape {
  static constexpr int FB = D == 2 ? 4 : 2, PB = 1, WPN = 4, WPP = 2;
};
template <int D, int ACT> struct RowAsyncShape<PlanHiddenFwd<D, ACT>> : WideHiddenShape<D> {};
// 2D hidden layers on split operands: 8 waves on a 128 x 32 tile (32
// features x 16 points each), three 40 KB slots
template <int ACT> struct RowAsyncShape<PlanHiddenFwdSplit<2, ACT>> {
  static constexpr int FB = 2, PB = 1, WPN = 4, WPP = 2;
};
// on first-order jets (six input planes): 8 waves on a 128 x 64 tile (32
// features x 32 points each), three 40 KB slots (per-mode 2D blocks: a
// 4,096-wide layer over 4,096 points 172 -> 155 us against 128 x 32 tiles)
template <int ACT> struct RowAsyncShape<PlanHiddenFwdSplit<JET1_2D, ACT>> {
  static constexpr int FB = 2, PB = 2, WPN = 4, WPP = 2;
};
// and the first-order reverse on 4-wave 64 x 32 tiles, four 10 KB slots
// (15.9 -> 15.6 ms per 2D per-mode step against the 8-wave 128 x 32 tile)
template <int ACT> struct RowAsyncShape<PlanHiddenBwd<JET1_2D, ACT>> {
  static constexpr int FB = 2, PB = 1, WPN = 2, WPP = 2;
};
// 2D output layer on split operands: 8 waves on a 256 x 32 tile (64 features
// x 16 points each), three 48 KB slots (four waves of the plan's own 128 x 32
// tile would leave two); each point block is read by 4 feature tiles, not 8
template <> struct RowAsyncShape<PlanOutFields2Split> {
  static constexpr int FB = 4, PB = 1, WPN = 4, WPP = 2;
};
template <int D, int ACT> struct RowAsyncShape<PlanHiddenBwd<D, ACT>> : WideHiddenShape<D> {};
// split GEMM of the factored reverse (65,536 x 512 x 512 per field group):
// 8 waves on a 256 x 64 tile, three 40 KB slots (0.210 -> 0.200 ms against
// the plan's 4-wave 128 x 64 tile; 128 x 128 and 256 x 32 tiles were slower)
template <> struct RowAsyncShape<PlanGemmSplit> {
  static constexpr int FB = 4, PB = 2, WPN = 4, WPP = 2;
};

// First term of a plan that reads input plane c (1 << 20: none).
template <class Plan>
__host__ __device__ constexpr int plan_first_use(int c) {
  for (int t = 0; t < Plan::NT; ++t)
    if (Plan::tc[t] == c) return t;
  return 1 << 20;
}

template <class Plan>
struct RowAsyncPol {
  using SH = RowAsyncShape<Plan>;
  static constexpr int RK = 32, LDK = 32;
  static constexpr int TN = SH::WPN * SH::FB * 16, TP = SH::WPP * SH::PB * 16;
  static constexpr int WT = TN * LDK, XT = TP * LDK;
  static constexpr int STAGE = Plan::NW * WT + Plan::CI * XT;  // bf16 elements per slot
  static constexpr int WPIECE = Plan::NW * TN / 16, XPIECE = Plan::CI * TP / 16;
  static constexpr int NPIECE = WPIECE + XPIECE;  // 1 KB copies per slot
  static constexpr int NWV = SH::WPN * SH::WPP, NTH = 64 * NWV;  // waves, threads per workgroup
  static constexpr int PPW = (NPIECE + NWV - 1) / NWV;  // per wave (the last round padded with repeats)
  static constexpr int SLOT_BYTES = STAGE * 2;
  // ring depth: as many slots as fit two 4-wave workgroups per CU (<= 80 KB
  // each) or one 8-wave workgroup (<= 144 KB)
  static constexpr int BUDGET = NWV == 8 ? 144 * 1024 : 80 * 1024;
  static constexpr int NS = SLOT_BYTES * 4 <= BUDGET ? 4 : (SLOT_BYTES * 3 <= BUDGET ? 3 : 2);
};

template <class Plan>
__global__ __launch_bounds__(RowAsyncPol<Plan>::NTH) __attribute__((amdgpu_waves_per_eu(2))) void rowgemm_async_kernel(
    RowArgs a, int64_t per_wg) {
  constexpr int P = PREC_BF16;
  using PR = Prec<P>;
  using S = __bf16;
  using M = float;
  using frag_t = typename PR::frag_t;
  using acc_t = typename PR::acc_t;
  using Q = RowAsyncPol<Plan>;
  constexpr int RK = Q::RK, LDK = Q::LDK, KE = PR::KE;
  constexpr int CI = Plan::CI, NW = Plan::NW, NA = Plan::NA, NT = Plan::NT;
  constexpr int FB = Q::SH::FB, PB = Q::SH::PB, WPN = Q::SH::WPN, WPP = Q::SH::WPP;
  constexpr int TN = Q::TN, TP = Q::TP, WT = Q::WT, XT = Q::XT, STAGE = Q::STAGE;
  constexpr int NS = Q::NS, NPIECE = Q::NPIECE, PPW = Q::PPW, WPIECE = Q::WPIECE;
  constexpr int NWV = Q::NWV, NTH = Q::NTH;
  static_assert(NWV == 4 || NWV == 8, "four or eight waves per workgroup");
  static_assert(NPIECE >= NWV, "a slot has a piece per wave");
  static_assert(FB % 2 == 0, "8-feature epilogue groups pair feature blocks");
  static_assert(NS >= 2 && NS <= 4, "ring of 2..4 slots");
  static_assert(PPW * (NS - 2) <= 63, "vmcnt range");

  // LDS pointers stay in address space 3 from the arrays on (a generic
  // pointer cast back to LDS would need a null check per use)
  typedef __attribute__((address_space(3))) S lds_S;
  typedef __attribute__((address_space(3))) const frag_t lds_frag;
  __shared__ __align__(16) S ring0[STAGE];
  __shared__ __align__(16) S ring1[STAGE];
  __shared__ __align__(16) S ring2[NS > 2 ? STAGE : 8];
  __shared__ __align__(16) S ring3[NS > 3 ? STAGE : 8];
  lds_S* const r0 = (lds_S*)ring0;
  lds_S* const r1 = (lds_S*)ring1;
  lds_S* const r2 = (lds_S*)ring2;
  lds_S* const r3 = (lds_S*)ring3;
  lds_S* const slots[4] = {r0, r1, r2, r3};

  const int64_t nfb = (a.nout + TN - 1) / TN, npb = (a.npts + TP - 1) / TP;
  const int64_t ntiles = npb * nfb;
  const int64_t t_beg = (int64_t)blockIdx.x * per_wg;
  const int64_t t_end = t_beg + per_wg < ntiles ? t_beg + per_wg : ntiles;
  // contraction length per tile: all of k, or with block-diagonal weights
  // (a.kwin; the launcher sets it only when every tile's window is whole and
  // of one length) max(TN, kwin) starting at the tile's first block
  const int64_t wlen = a.kwin > 0 ? (TN >= a.kwin ? (int64_t)TN : a.kwin) : a.k;
  const int nk = (int)((wlen + RK - 1) / RK);
  const int kpart = (int)(wlen - (int64_t)(nk - 1) * RK);  // valid k of the last k-step

  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wn = wave % WPN, wp = wave / WPN;
  const int fr = lane & 15, kg = (lane >> 4) * KE;
  constexpr int GE = Plan::template out_bytes<P>() == 2 ? 8 : 4;
  const int rg = PR::rowg(fr), rr = PR::rowr(fr);

  // ---- copies: this wave issues pieces q = wave + NWV t (t < PPW) of every
  // slot (a piece past NPIECE repeats an earlier one).  The per-wave plan is
  // set once; each issue is then straight-line code of exactly PPW copies,
  // so the compiler knows how many copies follow any load it must wait for.
  const int cr = lane >> 2, cu = lane & 3;   // row in the 16-row piece, LDS chunk slot
  const int cch = cu ^ ((cr >> 2) & 3);      // global chunk held by that slot (lds_k<S, 2>)
  const S* csrc[PPW];
  int64_t cld[PPW];
  int crow[PPW], cdst[PPW];
  bool cw[PPW];
#pragma unroll
  for (int t = 0; t < PPW; ++t) {
    const int q = wave + NWV * t;
    const int qq = q < NPIECE ? q : q - NPIECE;
    cw[t] = qq < WPIECE;
    if (cw[t]) {
      const int s = qq / (TN / 16), rb = qq % (TN / 16);
      csrc[t] = reinterpret_cast<const S*>(a.w[s]);
      cld[t] = a.ldw[s];
      crow[t] = 16 * rb;
      cdst[t] = s * WT + 16 * rb * LDK;
    } else {
      const int c = (qq - WPIECE) / (TP / 16), rb = (qq - WPIECE) % (TP / 16);
      csrc[t] = reinterpret_cast<const S*>(a.x[c]);
      cld[t] = a.ldx[c];
      crow[t] = 16 * rb;
      cdst[t] = NW * WT + c * XT + 16 * rb * LDK;
    }
  }
  // Copies past the reduction length re-read the row's last 16-byte chunk
  // inside it (k rounded up to 8 <= every row stride, the weight slices of a
  // lift included, whose stride spans several slices); those LDS columns
  // are zeroed before use.
  const int64_t kcap = (a.k + 7) / 8 * 8;
  auto issue = [&](int64_t tp, int64_t tf, int kt, lds_S* slot) __attribute__((always_inline)) {
    const int64_t p0 = tp * TP, f0 = tf * TN;
    const int64_t k0 = (a.kwin > 0 ? f0 / a.kwin * a.kwin : 0) + (int64_t)kt * RK;
#pragma unroll
    for (int t = 0; t < PPW; ++t) {
      const int64_t nrows = cw[t] ? a.nout : a.npts;
      int64_t row = (cw[t] ? f0 : p0) + crow[t] + cr;
      row = row < nrows ? row : nrows - 1;
      int64_t kk = k0 + 8 * cch;
      kk = kk + 8 <= kcap ? kk : kcap - 8;
      const S* src = csrc[t] + row * cld[t] + kk;
      lds_S* dst = slot + cdst[t];
      __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
    }
  };

  // Tiles are numbered feature-block-major (t = tf * npb + tp) and dealt in
  // contiguous ranges: the workgroups that read the same point block at the
  // same time are npb / per_wg apart, i.e. on one XCD (a multiple of 8 in the
  // launches here), so each point block comes from HBM once and each
  // workgroup's weight slice stays in L2.
  //
  // Every step issues one slot of copies, also past the end of the range
  // (the cursor then stays on the first tile past it: rows and k are clamped
  // in issue(), and the slot it fills is never read), so each wave always
  // has NS-2 slots issued after the one it waits for: one constant wait, and
  // the same count of copies on every path through the loop for the
  // compiler's own waits.  The kernel drains them before it ends.
  int64_t lt = t_beg;  // copy cursor (tile, its point / feature block, k-step)
  int64_t ltp = t_beg % npb, ltf = t_beg / npb;
  int lk = 0;
  auto advance = [&]() __attribute__((always_inline)) {
    if (lt >= t_end) return;
    if (++lk == nk) {
      lk = 0;
      ++lt;
      if (++ltp == npb) { ltp = 0; ++ltf; }
    }
  };
#pragma unroll
  for (int s = 0; s < NS - 1; ++s) {
    issue(ltp, ltf, lk, slots[s]);
    advance();
  }

  acc_t acc[NA][FB][PB];
  auto zero_acc = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int s = 0; s < NA; ++s)
#pragma unroll
      for (int i = 0; i < FB; ++i)
#pragma unroll
        for (int j = 0; j < PB; ++j) acc[s][i][j] = acc_t{0, 0, 0, 0};
  };
  zero_acc();

  // epilogue groups: lane (fr, g) holds points p0 + wp*PB*16 + 16j + fr and
  // features f0 + wn*FB*16 + (GE == 8 ? 32q + 8g : 16q + 4g) + [0, GE)
  constexpr int NQ = FB * 4 / GE;
  const int g = lane >> 4;
  auto group_p = [&](int64_t p0, int j) __attribute__((always_inline)) { return p0 + wp * PB * 16 + j * 16 + fr; };
  auto group_f = [&](int64_t f0, int q) __attribute__((always_inline)) {
    return f0 + wn * FB * 16 + (GE == 8 ? 32 * q + 8 * g : 16 * q + 4 * g);
  };
  typedef S tape_t __attribute__((ext_vector_type(GE)));

  // bias of this lane's epilogue features for the first two feature blocks,
  // in registers from the start (no load in the epilogue); more feature
  // blocks read it in the epilogue
  constexpr bool BIAS = PlanBiasCh0<Plan>::value;
  M breg[BIAS ? 2 : 1][BIAS ? NQ : 1][GE];
  const bool bias_reg = BIAS && nfb <= 2;
  if constexpr (BIAS) {
    const M* bsrc = reinterpret_cast<const M*>(a.bias);
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int e = 0; e < GE; ++e) {
          const int64_t f = group_f((int64_t)h * TN, q) + e;
          breg[h][q][e] = (bsrc && f < a.nout) ? bsrc[f] : (M)0;
        }
  }

  // Offsets of the operand fragment reads in a slot, opaque to the
  // compiler: with the channel stride folded into constant offsets, the reads
  // of channels past the first lose the LDS alias scope that tells the
  // compiler they do not touch the slot being refilled, and it then waits
  // for every copy in flight before them.
  int xoff[CI][PB];
#pragma unroll
  for (int j = 0; j < PB; ++j) {
    const int row = wp * PB * 16 + j * 16 + fr;
#pragma unroll
    for (int c = 0; c < CI; ++c) {
      xoff[c][j] = NW * WT + c * XT + row * LDK + lds_k<S, 2>(row, kg);
      opaque_vgpr(xoff[c][j]);
    }
  }

  int64_t ct = t_beg;  // compute cursor
  int64_t ctp = t_beg % npb, ctf = t_beg / npb;
  int ck = 0;
  // always inlined: the accumulators must stay in registers
  constexpr int TC = PlanTapeCh<Plan>::value;
  auto step = [&](lds_S* cur, lds_S* refill) __attribute__((always_inline)) -> bool {
    if (ct >= t_end) return false;
    __builtin_amdgcn_s_waitcnt(vmcnt_imm(PPW * (NS - 2)));
    ring_barrier();
    if (ck == nk - 1 && kpart < RK) {  // zero k >= the reduction length
      constexpr int ROWS = NW * TN + CI * TP;
      const int nz = RK - kpart;
      for (int e = threadIdx.x; e < ROWS * nz; e += NTH) {
        const int row = e / nz, k = kpart + e % nz;
        cur[row * LDK + lds_k<S, 2>(row, k)] = (S)0.0f;
      }
      ring_barrier();
    }
    tape_t traw[PB][NQ][TC > 0 ? TC : 1];
    bool tape_now = false;
    if constexpr (TC > 0) {
      tape_now = ck == nk - 1 && (ctp * TP + TP <= a.npts) && (ctf * TN + TN <= a.nout);
      if (tape_now) {
        const S* tz = reinterpret_cast<const S*>(a.tape);
#pragma unroll
        for (int j = 0; j < PB; ++j)
#pragma unroll
          for (int q = 0; q < NQ; ++q)
#pragma unroll
            for (int c = 0; c < TC; ++c)
              traw[j][q][c] = *reinterpret_cast<const tape_t*>(
                  tz + group_p(ctp * TP, j) * a.ldt + group_f(ctf * TN, q) + c * a.tstride);
      }
    }
    issue(ltp, ltf, lk, refill);
    advance();
    const lds_S* ws = cur;
    frag_t wf[NW][FB], xf[CI][PB];
#pragma unroll
    for (int i = 0; i < FB; ++i) {
      const int row = wn * FB * 16 + (GE == 8 ? 32 * (i >> 1) + 8 * rg + 4 * (i & 1) : 16 * i + 4 * rg) + rr;
      const int off = row * LDK + lds_k<S, 2>(row, kg);
#pragma unroll
      for (int s = 0; s < NW; ++s) wf[s][i] = *reinterpret_cast<lds_frag*>(ws + s * WT + off);
    }
    // The terms run in chunks of TG (>= 12 MFMAs each); an input plane is
    // read just after the first MFMA of the chunk before the one that first
    // uses it, so its LDS latency runs under that chunk's MFMAs.  (Read all
    // at once, the compiler reused a few registers for the planes in turn
    // and waited lgkmcnt(0) every two MFMAs: the wait also covers every
    // read issued before it, so a read cannot overlap the MFMAs before its
    // own wait unless the MFMAs come between.)
    constexpr int TG = (12 + FB * PB - 1) / (FB * PB), NCH = (NT + TG - 1) / TG;
    auto read_chunk = [&](int k) __attribute__((always_inline)) {
#pragma unroll
      for (int c = 0; c < CI; ++c)
        if (plan_first_use<Plan>(c) / TG == k)
#pragma unroll
          for (int j = 0; j < PB; ++j) xf[c][j] = *reinterpret_cast<lds_frag*>(cur + xoff[c][j]);
    };
    auto term = [&](int t, int i, int j) __attribute__((always_inline)) {
      const frag_t wv = Plan::tneg[t] ? fneg(wf[Plan::tw[t]][i]) : wf[Plan::tw[t]][i];
      acc[Plan::ta[t]][i][j] = PR::mfma(wv, xf[Plan::tc[t]][j], acc[Plan::ta[t]][i][j]);
    };
    read_chunk(0);
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int t0 = k * TG, t1 = t0 + TG < NT ? t0 + TG : NT;
      term(t0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (k + 1 < NCH) read_chunk(k + 1);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = t0; t < t1; ++t)
#pragma unroll
        for (int i = 0; i < FB; ++i)
#pragma unroll
          for (int j = 0; j < PB; ++j)
            if (t != t0 || i != 0 || j != 0) term(t, i, j);
    }
    if (++ck == nk) {
      const int64_t p0 = ctp * TP, f0 = ctf * TN;
      auto group_v = [&](int j, int q, M (*v)[GE]) __attribute__((always_inline)) {
#pragma unroll
        for (int s = 0; s < NA; ++s)
#pragma unroll
          for (int e = 0; e < GE; ++e)
            v[s][e] = GE == 8 ? (M)acc[s][2 * q + (e >> 2)][j][e & 3] : (M)acc[s][q][j][e];
      };
      if constexpr (TC > 0) {
        if (tape_now) {
#pragma unroll
          for (int j = 0; j < PB; ++j)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
              M v[NA][GE], u[TC][GE];
              group_v(j, q, v);
#pragma unroll
              for (int c = 0; c < TC; ++c)
#pragma unroll
                for (int e = 0; e < GE; ++e) u[c][e] = tomath<M>(traw[j][q][c][e]);
              Plan::template epilogue_tape<P, GE>(a, group_p(p0, j), group_f(f0, q), GE, v, u);
            }
        }
      }
      if (!tape_now) {
#pragma unroll
        for (int j = 0; j < PB; ++j) {
          const int64_t p = group_p(p0, j);
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const int64_t f = group_f(f0, q);
            const int64_t rem = a.nout - f;
            if (p < a.npts && rem > 0) {
              const int nvalid = rem >= GE ? GE : (int)rem;
              M v[NA][GE];
              group_v(j, q, v);
              if constexpr (BIAS) {
                if (bias_reg) {
                  const M w1 = (M)(ctf != 0), w0 = (M)1 - w1;
#pragma unroll
                  for (int e = 0; e < GE; ++e) v[0][e] += w0 * breg[0][q][e] + w1 * breg[1][q][e];
                } else {
                  const M* bsrc = reinterpret_cast<const M*>(a.bias);
#pragma unroll
                  for (int e = 0; e < GE; ++e) v[0][e] += (bsrc && e < nvalid) ? bsrc[f + e] : (M)0;
                }
                Plan::template epilogue<P, GE, M, true>(a, p, f, nvalid, v);
              } else {
                Plan::template epilogue<P, GE>(a, p, f, nvalid, v);
              }
            }
          }
        }
      }
      zero_acc();
      ck = 0;
      ++ct;
      if (++ctp == npb) { ctp = 0; ++ctf; }
    }
    return true;
  };
  if constexpr (NS == 4) {
    while (step(r0, r3) && step(r1, r0) && step(r2, r1) && step(r3, r2)) {
    }
  } else if constexpr (NS == 3) {
    while (step(r0, r2) && step(r1, r0) && step(r2, r1)) {
    }
  } else {
    while (step(r0, r1) && step(r1, r0)) {
    }
  }
  __builtin_amdgcn_s_waitcnt(vmcnt_imm(0));  // copies past the range land before the LDS is released
}

// Plans whose slot leaves a ring of fewer than 3 slots stay on the
// register-staged rowgemm_kernel, whose two register stages in flight did
// better there (measured on the 2D output-layer reverse; halving its point
// tile to fit 3 slots doubled the weight traffic per point and was slower
// still).
// The 3D split hidden layers (20 input planes: 36 KB slots) run on a
// two-slot ring: one k-step in flight, like the register-staged kernel, but
// the persistent tile walk issues the next tile's copies during an epilogue
// (a 131,072-point 3D layer 5.0 -> 4.1 ms; per-mode blocks with 4 k-steps
// a tile, 186 -> 167 us).
template <class Plan> struct RowAsyncTwoSlots : std::false_type {};
template <int ACT> struct RowAsyncTwoSlots<PlanHiddenFwdSplit<3, ACT>> : std::true_type {};
template <class Plan>
constexpr bool row_async_ok() {
  return RowAsyncPol<Plan>::NS >= 3 || (RowAsyncPol<Plan>::NS == 2 && RowAsyncTwoSlots<Plan>::value);
}
