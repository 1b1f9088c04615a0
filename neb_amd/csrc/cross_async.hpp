// Split-R cross GEMM for operands stored in bf16 (NL = 1: one plane, the
// CR_BF16 product; NL = 2: hi and lo planes, the CR_BF16X3 product
// hi.hi + hi.lo + lo.hi), fed by asynchronous global -> LDS copies.
//
// Same contraction, partial-tile layout and reduction as cross_kernel (with
// fp32 partial tiles)
// (crossgemm.hpp); what differs is the data path:
//   * global_load_lds_dwordx4 moves 16 bytes per lane straight into LDS, so
//     no VGPRs stage the tiles and no VALU converts them; a ring of NS LDS
//     stages keeps NS-1 k-steps of loads in flight per workgroup;
//   * tiles stay k-major in LDS (the row-major source layout, 32 rows x 128
//     columns per plane) and the MFMA fragments (8 consecutive k per lane)
//     are read with ds_read_b64_tr_b16: within each 16-lane group, lane i
//     receives element i % 4 of the 8-byte blocks addressed by lanes
//     i/4 + 4e (e = 0..3).  Pointing lane m at row 8kg + m/4 (+4 for the
//     second read), columns c + 4 (m % 4) .. +3 hands lane i column c + i,
//     rows 8kg .. 8kg+3 (then +4 .. +7): the bf16 16x16x32 MFMA operand.
//   * 16-byte column units of k-row r sit at unit u ^ swz(r),
//     swz(r) = 2 (r % 4) + 8 ((r / 8) % 2), which spreads the 32 units one
//     transposing read touches over all 16 bank groups twice (the 512-byte
//     minimum of two cycles).
// WM x WN waves each own a (128/WM) x (128/WN) block of the 128 x 128 tile.
// Rows past a segment's end are zeroed in LDS before use; columns past the
// matrix (inside the padded rows) only reach output rows / columns that are
// never stored.  Segment signs must be +-1 (applied by flipping the sign
// bits of the X fragments); per-row weights are not supported here.
#pragma once
#include "crossgemm.hpp"

// NL planes per operand, WM x WN waves over the TT x TT tile (each wave
// (TT / WM) x (TT / WN)), NS ring slots.
template <int NL, int WM, int WN, int NS_>
struct AsyncPol {
  static constexpr int TT = CR_TM;             // output tile side
  static constexpr int NS = NS_;               // LDS ring slots
  static constexpr int NWAVE = WM * WN;
  static constexpr int BI = TT / WM / 16, BJ = TT / WN / 16;  // 16x16 blocks per wave
  static constexpr int RK = 32;                // k-rows per stage
  static constexpr int PLANE = RK * TT;        // bf16 elements per plane tile
  static constexpr int STAGE = 2 * NL * PLANE; // X planes then Y planes
  static constexpr int UPR = TT / 8;           // 16-byte units per k-row
  static constexpr int RPP = 64 / UPR;         // k-rows per 1 KB copy piece
  static constexpr int PPP = RK / RPP;         // pieces per plane tile
  static constexpr int LPW = 2 * NL * PPP / NWAVE;  // 1 KB copies per wave per stage
  static constexpr int FLUSH = NL == 2 ? 256 : (1 << 30);
  static_assert(2 * NL * PPP % NWAVE == 0, "copy pieces split evenly over the waves");
};

// Work item of workgroup L: the dispatcher places workgroup L on XCD L % 8,
// and the N items are dealt to the XCDs in contiguous blocks (XCD x holds the
// ceil((N - x) / 8) workgroups L = x, x + 8, ...), so the tiles one XCD runs
// at a time share their row splits, which then come from HBM once per XCD.
__device__ inline int async_item(int L, int N) {
  const int x = L & 7, y = L >> 3, base = N >> 3, rem = N & 7;
  return x * base + (x < rem ? x : rem) + y;
}

__device__ inline int async_swz(int kr) { return 2 * (kr & 3) + 8 * ((kr >> 3) & 1); }

__device__ inline bf16x8 async_neg(bf16x8 v) {  // exact negation: flip the sign bits
  typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
  u16x8 u = __builtin_bit_cast(u16x8, v);
  u ^= (unsigned short)0x8000;
  return __builtin_bit_cast(bf16x8, u);
}

// k-step cursor over the (segment, row) sequence of one row split
struct AsyncCursor {
  int s;
  int64_t r0, rend;
  __device__ inline void seek(const CrossArgs& a, int split, int from) {
    for (s = from; s < a.nseg; ++s) {
      int64_t rb, re;
      split_range<32>(a.seg[s].rows, a.splits, split, rb, re);
      if (rb < re) { r0 = rb; rend = re; return; }
    }
  }
  __device__ inline void next(const CrossArgs& a, int split) {
    r0 += 32;
    if (r0 >= rend) seek(a, split, s + 1);
  }
};

template <int NL, int WM, int WN, int NS_, bool NEG>
__global__ __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(NL == 1 ? 4 : 2))) void
cross_async_kernel(CrossArgs a) {
  using Pol = AsyncPol<NL, WM, WN, NS_>;
  constexpr int TT = Pol::TT;
  constexpr int NS = Pol::NS, RK = Pol::RK, PLANE = Pol::PLANE, STAGE = Pol::STAGE, LPW = Pol::LPW;
  constexpr int BI = Pol::BI, BJ = Pol::BJ, NT = 64 * Pol::NWAVE;
  constexpr int UPR = Pol::UPR, RPP = Pol::RPP, PPP = Pol::PPP;
  typedef short v4i16 __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(3))) v4i16 lds_v4;
  typedef __attribute__((address_space(3))) void lds_void;
  // one LDS object per ring slot: the compiler's wait insertion tracks the
  // asynchronous copies per object, so a read of slot k waits for slot k's
  // copies only and the copies into the other slots stay in flight
  static_assert(NS >= 2 && NS <= 4, "the k-loop below is unrolled over a ring of 2..4 slots");
  __shared__ __align__(16) __bf16 ring0[STAGE];
  __shared__ __align__(16) __bf16 ring1[STAGE];
  __shared__ __align__(16) __bf16 ring2[NS > 2 ? STAGE : 8];
  __shared__ __align__(16) __bf16 ring3[NS > 3 ? STAGE : 8];

  int split, tile;
  {
    const int L = blockIdx.x;
    const int N = a.ntiles * a.splits;
    const int u = async_item(L, N);
    split = u / a.ntiles;
    tile = u % a.ntiles;
  }
  int ti, tj;
  cross_tile_ij(a, tile, ti, tj);
  const int64_t i0 = (int64_t)ti * TT, j0 = (int64_t)tj * TT;
  // wave index made provably uniform, so the segment descriptors behind the
  // per-wave copy pieces are read with scalar loads (a vector load would be
  // counted by vmcnt, and waiting for it would also wait for every copy)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = wave / WN, wn = wave % WN;  // this wave's (16 BI) x (16 BJ) block of the tile

  // ---- load side: this wave's 1 KB pieces of a stage -----------------------
  // piece q = wave * LPW + t: operand op = q / (PPP NL), plane q / PPP % NL,
  // k-rows RPP (q % PPP) .. +RPP-1 (RPP = 4 rows of UPR = 16 units).  Lane
  // p fills LDS unit p of the piece: k-row RPP (q % PPP) + p / UPR, unit
  // p % UPR, holding global column unit (p % UPR) ^ swz(row).
  // A full k-step loads from the uniform address of its first row plus a
  // per-lane byte offset that is fixed for the whole segment (set when the
  // load cursor enters it); only a segment's last, partial k-step clamps
  // rows (and pays the 64-bit address arithmetic).  Offsets fit 32 bits:
  // the host keeps 32 ld bytes below 2^31.
  const int prow = lane / UPR, punit = lane % UPR;
  // per piece, fixed for the kernel: LDS destination (bf16 elements from the
  // slot), k-row; per segment (set when the load cursor enters one): source
  // plane, row stride and the lane's byte offset -- so the steady-state
  // k-step reads no segment descriptor (a scalar load left in flight there
  // would turn every LDS wait of the step into lgkmcnt(0))
  int pdst[LPW], pkr[LPW], pop[LPW];
#pragma unroll
  for (int t = 0; t < LPW; ++t) {
    const int q = wave * LPW + t;
    pop[t] = q / (PPP * NL);
    const int pl = (q / PPP) % NL, kb = q % PPP;
    pdst[t] = (pop[t] * NL + pl) * PLANE + kb * RPP * TT;
    pkr[t] = RPP * kb + prow;
  }
  int lseg = -1;
  uint32_t poff[LPW];
  const __bf16* psrc[LPW];
  int64_t pld[LPW];
  auto issue = [&](const AsyncCursor& c, __bf16* slot) __attribute__((always_inline)) {
    if (c.s != lseg) {
      const CrossSeg& sg = a.seg[c.s];
#pragma unroll
      for (int t = 0; t < LPW; ++t) {
        const int q = wave * LPW + t;
        const int op = pop[t], pl = (q / PPP) % NL;
        const int64_t ld = op == 0 ? sg.ldx : sg.ldy;
        psrc[t] = reinterpret_cast<const __bf16*>(op == 0 ? (pl == 0 ? sg.x : sg.xlo) : (pl == 0 ? sg.y : sg.ylo));
        pld[t] = ld;
        int64_t gc = (op == 0 ? i0 : j0) + 8 * (punit ^ async_swz(pkr[t]));
        gc = gc + 8 <= ld ? gc : ld - 8;
        poff[t] = (uint32_t)((pkr[t] * ld + gc) * 2);
      }
      lseg = c.s;
    }
    // one branch for the whole step (per-piece branches kept the clamped
    // path's 64-bit state live through the full one)
    if (c.r0 + RK <= c.rend) {
#pragma unroll
      for (int t = 0; t < LPW; ++t) {
        const char* base = reinterpret_cast<const char*>(psrc[t] + c.r0 * pld[t]);
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const __bf16*>(base + poff[t]), (lds_void*)(slot + pdst[t]),
                                         16, 0, 0);
      }
    } else {
#pragma unroll
      for (int t = 0; t < LPW; ++t) {
        const int64_t ld = pld[t];
        int64_t gr = c.r0 + pkr[t];
        gr = gr < c.rend ? gr : c.rend - 1;
        int64_t gc = (pop[t] == 0 ? i0 : j0) + 8 * (punit ^ async_swz(pkr[t]));
        gc = gc + 8 <= ld ? gc : ld - 8;
        // (the address through a named pointer: passed as the expression
        // itself, the host pass dropped the kernel's launch stub)
        const __bf16* gsrc = psrc[t] + gr * ld + gc;
        __builtin_amdgcn_global_load_lds(gsrc, (lds_void*)(slot + pdst[t]), 16, 0, 0);
      }
    }
  };

  // ---- compute side: per-lane byte offsets of the transposing reads --------
  // lane = 16 kg + m reads k-row 8 kg + m / 4 (+4 for the second half),
  // columns cb + 4 (m % 4): unit cb / 8 + (m % 4) / 2, half (m % 2).
  const int kg = lane >> 4, m = lane & 15;
  const int kr0 = 8 * kg + (m >> 2);
  const int sw = async_swz(kr0);  // equal for kr0 + 4
  auto rd_off = [&](int cb) __attribute__((always_inline)) {
    const int unit = (cb >> 3) + ((m & 3) >> 1);
    return kr0 * (TT * 2) + 16 * (unit ^ sw) + 8 * (m & 1);  // bytes within a plane tile
  };
  int xoff[BI], yoff[BJ];
#pragma unroll
  for (int i = 0; i < BI; ++i) xoff[i] = rd_off(wm * 16 * BI + 16 * i);
#pragma unroll
  for (int j = 0; j < BJ; ++j) yoff[j] = rd_off(wn * 16 * BJ + 16 * j);

  auto frag = [&](const __bf16* slot, int byte_off) __attribute__((always_inline)) -> bf16x8 {
    typedef __attribute__((address_space(3))) unsigned char lds_u8;
    lds_v4* p0 = (lds_v4*)((lds_u8*)(slot) + byte_off);
    lds_v4* p1 = (lds_v4*)((lds_u8*)(slot) + byte_off + 4 * TT * 2);
    const v4i16 lo4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16(p0);
    const v4i16 hi4 = __builtin_amdgcn_ds_read_tr16_b64_v4i16(p1);
    typedef short v8i16 __attribute__((ext_vector_type(8)));
    const v8i16 v = {lo4[0], lo4[1], lo4[2], lo4[3], hi4[0], hi4[1], hi4[2], hi4[3]};
    return __builtin_bit_cast(bf16x8, v);
  };

  constexpr bool TWO = Pol::FLUSH < (1 << 20);
  f32x4 inner[BI][BJ];
  float outer[TWO ? BI : 1][TWO ? BJ : 1][TWO ? 4 : 1];
#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j) inner[i][j] = f32x4{0, 0, 0, 0};
  if constexpr (TWO) {
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) outer[i][j][r] = 0.f;
  }

  // ---- pipeline ------------------------------------------------------------
  AsyncCursor cl, cc;  // load cursor (runs NS-1 steps ahead), compute cursor
  cl.seek(a, split, 0);
  cc = cl;
  int nissued = 0, k = 0, since = 0;
  __bf16* const slots[4] = {ring0, ring1, ring2, ring3};
#pragma unroll
  for (int t = 0; t < NS - 1; ++t)
    if (cl.s < a.nseg) {
      issue(cl, slots[t]);
      ++nissued;
      cl.next(a, split);
    }
  // fp32 partial tile: the sums are fp32 already (exactly the values an FP64
  // tile would hold, at half the bytes); the reduction adds in FP64.  Lane
  // (kg, m) holds rows 4 kg + r, column m of each 16 x 16 block.
  float* const out = reinterpret_cast<float*>(a.partial) + ((int64_t)split * a.ntiles + tile) * (TT * TT);
  auto tile_idx = [&](int i, int j, int r) __attribute__((always_inline)) {
    const int li = wm * 16 * BI + i * 16 + 4 * (lane >> 4) + r;
    const int lj = wn * 16 * BJ + j * 16 + (lane & 15);
    return li * TT + lj;
  };
  // step k computes from slot k % NS and refills slot (k - 1) % NS
  auto step = [&](__bf16* cur, __bf16* refill) __attribute__((always_inline)) -> bool {
    if (cc.s >= a.nseg) return false;
    // this wave's copies of step k are done once at most (nissued - k - 1)
    // later steps' copies are outstanding; the barrier then covers every
    // wave's, and every wave is past step k-1, whose slot is refilled below
    wait_copies<LPW, NS>(nissued - k - 1);
    ring_barrier();
    const int64_t valid = cc.rend - cc.r0;
    if (valid < RK) {  // last k-step of a segment: zero the rows past its end
      const int nz = (RK - (int)valid) * UPR;
      for (int e = threadIdx.x; e < 2 * NL * nz; e += NT) {
        const int p = e / nz, rem = e % nz;
        const int kr = (int)valid + rem / UPR, un = rem % UPR;
        *reinterpret_cast<uint4*>(cur + p * PLANE + kr * TT + 8 * un) = uint4{0, 0, 0, 0};
      }
      ring_barrier();
    }
    if (cl.s < a.nseg) {
      issue(cl, refill);
      ++nissued;
      cl.next(a, split);
    }
    bf16x8 xh[BI], xl[BI];
#pragma unroll
    for (int i = 0; i < BI; ++i) {
      xh[i] = frag(cur, xoff[i]);
      if constexpr (NL == 2) xl[i] = frag(cur, PLANE * 2 + xoff[i]);
    }
    if constexpr (NEG) {  // some segment has sign -1 (uniform per k-step)
      if (a.seg[cc.s].sign < 0) {
#pragma unroll
        for (int i = 0; i < BI; ++i) {
          xh[i] = async_neg(xh[i]);
          if constexpr (NL == 2) xl[i] = async_neg(xl[i]);
        }
      }
    }
    // Y fragments double-buffered in registers: block j + 1's transposing
    // reads are issued right after block j's first MFMA (whose operand wait
    // then covers block j's reads only) and the rest of block j's MFMAs run
    // while they are in flight.  With one register set the reads of block
    // j + 1 waited for block j's MFMAs and exposed their whole latency; read
    // before block j's MFMAs they fell under block j's own lgkmcnt(0) wait.
    bf16x8 ynh = frag(cur, NL * PLANE * 2 + yoff[0]);
    bf16x8 ynl = ynh;
    if constexpr (NL == 2) ynl = frag(cur, NL * PLANE * 2 + PLANE * 2 + yoff[0]);
#pragma unroll
    for (int j = 0; j < BJ; ++j) {
      const bf16x8 yh = ynh;
      const bf16x8 yl = ynl;
      if constexpr (NL == 2) inner[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[0], yh, inner[0][j], 0, 0, 0);
      else inner[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[0], yh, inner[0][j], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (j + 1 < BJ) {
        const int jn = j + 1 < BJ ? j + 1 : j;
        ynh = frag(cur, NL * PLANE * 2 + yoff[jn]);
        if constexpr (NL == 2) ynl = frag(cur, NL * PLANE * 2 + PLANE * 2 + yoff[jn]);
      }
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (NL == 2) {
        inner[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[0], yl, inner[0][j], 0, 0, 0);
        inner[0][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[0], yh, inner[0][j], 0, 0, 0);
#pragma unroll
        for (int i = 1; i < BI; ++i) {
          inner[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xl[i], yh, inner[i][j], 0, 0, 0);
          inner[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], yl, inner[i][j], 0, 0, 0);
          inner[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], yh, inner[i][j], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int i = 1; i < BI; ++i)
          inner[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xh[i], yh, inner[i][j], 0, 0, 0);
      }
    }
    cc.next(a, split);
    ++k;
    if constexpr (TWO) {
      since += RK;
      if (since >= Pol::FLUSH || cc.s >= a.nseg) {
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < BJ; ++j) {
#pragma unroll
            for (int r = 0; r < 4; ++r) outer[i][j][r] += inner[i][j][r];
            inner[i][j] = f32x4{0, 0, 0, 0};
          }
        since = 0;
      }
    }
    return true;
  };
  if constexpr (NS == 4) {
    while (step(ring0, ring3) && step(ring1, ring0) && step(ring2, ring1) && step(ring3, ring2)) {
    }
  } else if constexpr (NS == 3) {
    while (step(ring0, ring2) && step(ring1, ring0) && step(ring2, ring1)) {
    }
  } else {
    while (step(ring0, ring1) && step(ring1, ring0)) {
    }
  }

#pragma unroll
  for (int i = 0; i < BI; ++i)
#pragma unroll
    for (int j = 0; j < BJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if constexpr (TWO) out[tile_idx(i, j, r)] = outer[i][j][r];
        else out[tile_idx(i, j, r)] = inner[i][j][r];
      }
}
