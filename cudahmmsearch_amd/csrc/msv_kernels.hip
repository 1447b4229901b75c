// msv_kernels.hip -- gfx950 kernels for HMMER3's MSV filter.
//
// Three stages per (profile, target) pair, together bit-exact to the
// reference composite p7_MSVFilter (impl_sse/msvfilter.c:78-213):
//
//  1. msv_fast    J-free max-plus recurrence over all cells with the begin
//                 score B0 as a floor -- the quantity the reference's SSV path
//                 computes (impl_sse/ssvfilter.c:1364-1413) -- in packed f16
//                 (2 cells per VGPR, clamp bit = floor), followed by SSV's
//                 decision rules.  Pairs whose J state could fire are queued.
//  2. msv_slow    the byte-saturating MSV row loop with E/J/B updates and the
//                 255-overflow test, literally as impl_sse/msvfilter.c:112-210,
//                 in packed u16 lanes, for the queued pairs.
//  3. msv_finalize  usc and the F1 pass bit from host-built tables.
//
// Wavefront layout: a lane group of G lanes (8, 16, 32 or 64) owns one target;
// its 2G f16 slots hold model positions striped as k = j + z*Q + 1 (slot z,
// register j).  The diagonal dependency of the last register moves one slot
// per row through a DPP lane shift (wave_shr:1 / row_shr:1) plus v_alignbit.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>
#include "msv_device.h"
#include "msv_cells.h"

// ---------------------------------------------------------------------------
// Stage 1: fast J-free pass.
// ---------------------------------------------------------------------------

template <int G, int Q>
__device__ __forceinline__ void fast_task(const FastArgs &a, const ProfDev &pd, uint16_t *res, int64_t rstride,
                                          const uint4 *lds, int64_t s, int lane) {
  constexpr int QB = (Q + 3) / 4;
  const bool valid = s < a.db.n;
  const int L = valid ? a.db.len[s] : 0;
  const int Lpad = (L + 15) & ~15;
  const uint8_t *p = a.db.res + (valid ? a.db.off[s] : 0);

  int Lmax = Lpad;
  if constexpr (G < 64) {
#pragma unroll
    for (int o = 32; o >= G; o >>= 1) Lmax = max(Lmax, __shfl_xor(Lmax, o, 64));
  }
  Lmax = __builtin_amdgcn_readfirstlane(Lmax);

  const uint32_t nb = (uint32_t) a.db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);
  const uint4 *tl = lds + lane;

  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;
  uint32_t acc = 0u;

  // Residues arrive 16 at a time; the next group is loaded while the current
  // one is processed, and each row's emissions are read one row ahead (the
  // LDS read of row i+1 is in flight while row i computes).
  uint4 cnext = *reinterpret_cast<const uint4 *>(p);
  uint4 c = (0 < Lpad) ? cnext : null4;
  uint32_t e[QB * 4];
  load_row<Q>(tl, byte_of(c, 0), e);
  for (int i0 = 0; i0 < Lmax; i0 += 16) {
    const int i1 = i0 + 16;
    cnext = *reinterpret_cast<const uint4 *>(p + (i1 < Lpad ? i1 : 0));
    const uint4 cn = (i1 < Lpad) ? cnext : null4;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) {
      uint32_t en[QB * 4];
      load_row<Q>(tl, (rr < 15) ? byte_of(c, rr + 1) : byte_of(cn, 0), en);
      const uint32_t sh = shift_slot<G>(r[Q - 1], lane);
#pragma unroll
      for (int j = Q - 1; j >= 1; --j) r[j] = cell(r[j - 1], e[j]);
      r[0] = cell(sh, e[0]);
      int j = 0;
#pragma unroll
      for (; j + 1 < Q; j += 2) acc = max3h(acc, r[j], r[j + 1]);
      if constexpr ((Q & 1) != 0) acc = max2h(acc, r[Q - 1]);
#pragma unroll
      for (int k = 0; k < Q; ++k) e[k] = en[k];
      row_fence<Q>(r);
      asm volatile("" : "+v"(acc));
    }
    c = cn;
  }

  // Group maximum of both f16 halves, then the reference's decision.
  const h2v ah = as_h2(acc);
  float m = fmaxf((float) ah.x, (float) ah.y);
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (valid && (lane % G) == 0) {
    const int Rp = (int) (m * 2048.0f + 0.5f);
    const uint32_t w = decide(Rp, a.db.tjb[s], pd);
    res[s * rstride] = (uint16_t) w;
    if ((w >> 8) == RES_SLOW && a.slow_count)
      defer(w, a.prof_idx, s, a.slow, a.slow_count, a.slow_cap, a.jslow, a.jslow_count, a.jslow_cap);
  }
}

template <int G, int Q>
__global__ void __launch_bounds__(1024) msv_fast(FastArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int QB = (Q + 3) / 4;
  constexpr int NG = 64 / G;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;

  const int ntab = (a.db.R + 1) * QB * 64;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) lds[i] = a.prof.ftab[i];
  __syncthreads();

  const int g = lane / G;
  if (a.mode == 0) {
    for (;;) {
      int c = 0;
      if (lane == 0) c = atomicAdd(a.counter, 1);
      c = __builtin_amdgcn_readfirstlane(c);
      if (c >= a.nchunks) break;
      const int t1 = a.chunk_start[c + 1];
      for (int t = a.chunk_start[c]; t < t1; ++t) fast_task<G, Q>(a, a.prof, a.prof.res, 1, lds, (int64_t) t * NG + g, lane);
    }
  } else if (a.mode == 2) {
    // One chunk per workgroup, waves interleaved over its tasks: the launch
    // is not persistent, so workgroups of concurrent searches (one per
    // stream) share the CUs as they come and go.
    const int c = blockIdx.x;
    if (c < a.nchunks) {
      const int t1 = a.chunk_start[c + 1];
      for (int t = a.chunk_start[c] + wave; t < t1; t += nwaves)
        fast_task<G, Q>(a, a.prof, a.prof.res, 1, lds, (int64_t) t * NG + g, lane);
    }
  } else {
    for (int t = blockIdx.x * nwaves + wave; t < a.ntasks; t += gridDim.x * nwaves)
      fast_task<G, Q>(a, a.prof, a.prof.res, 1, lds, a.order[t], lane);
  }
}

// ---------------------------------------------------------------------------
// Stage 1b: the same J-free pass over one column segment of a long model.
// One target per wave (G = 64), static round-robin over the sorted targets.
// Row i's first position takes as its diagonal predecessor the value that
// left the previous segment's last position in row i-1 (the carry: one u16
// per residue, indexed like the residues, overwritten in place by this
// segment's own carry); the running maximum of the earlier segments arrives
// per target in partial[].
//
// The carries move by DPP, not through scalar registers.  One register x
// holds two queues: lanes 0..15 the incoming carries of a block of 16 rows
// (lane 0 = the current row's), lanes 48..63 the outgoing ones (lane 63 =
// the latest).  Each row one row_shl:1 advances both queues (the row's
// wave shift takes lane 0 of x where it leaves a hole, and lane 63, which
// has no source, takes the value that left the segment's last position in
// the row before); the block's outgoing carries are stored once per block.
// Residues come through the scalar cache, so a row's table address is
// scalar arithmetic plus one v_add; only a target's real rows are computed
// (its last block runs L mod 16 rows).
// ---------------------------------------------------------------------------

// One row of a segment: the descending register update with the carry
// entering lane 0's low half, and the row's cells into the running maximum
// (odd Q: an even row leaves register 0 to the odd row after it).
template <int Q, int RR>
__device__ __forceinline__ void seg_row(uint32_t (&r)[Q], const uint32_t *e, uint32_t &acc, uint32_t &x) {
  const uint32_t last = r[Q - 1];
  const uint32_t xs = __builtin_amdgcn_update_dpp(last, x, 0x101 /* row_shl:1 */, 0xf, 0xf, false);
  // wave_shr:1; lane 0 has no source and keeps x, whose high half there is
  // this row's carry (alignbit moves it into the low half of sh)
  const uint32_t prev = __builtin_amdgcn_update_dpp(x, last, 0x138, 0xf, 0xf, false);
  x = xs;
  const uint32_t sh = __builtin_amdgcn_alignbit(last, prev, 16);
  const uint32_t r0_prev = r[0];
#pragma unroll
  for (int j = Q - 1; j >= 1; --j) r[j] = cell(r[j - 1], e[j]);
  r[0] = cell(sh, e[0]);
  if constexpr ((Q & 1) != 0 && (RR & 1) != 0) acc = max3h(acc, r0_prev, r[Q - 1]);
  if constexpr ((Q & 1) != 0 && (RR & 1) == 0) {
#pragma unroll
    for (int j = 1; j + 1 < Q; j += 2) acc = max3h(acc, r[j], r[j + 1]);
  } else {
#pragma unroll
    for (int j = 0; j + 1 < Q; j += 2) acc = max3h(acc, r[j], r[j + 1]);
  }
}

template <int Q>
__device__ __forceinline__ void seg_task(const SegArgs &a, const uint4 *tl, int s, int lane) {
  constexpr int QB = (Q + 3) / 4;
  const int L = sload(a.db.len + s);
  const int Lpad = (L + 15) & ~15;
  const int Lfull = L & ~15;
  const int64_t base = sload(a.db.off + s);
  const uint8_t *p = a.db.res + base;
  uint16_t *cy = a.carry + base;
  const uint32_t nb = (uint32_t) a.db.R * 0x01010101u;

  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;
  uint32_t acc = 0u;
  uint32_t clast = 0u;               // carry of the row before the block, high half

  // Loads for a block are issued one block ahead, so their latency is
  // covered by the block before.  Residues: dword (lane & 3) of the 16-byte
  // block (the null row past the target's end), made wave-uniform with
  // readlane when the block starts.  Incoming carries: rows i .. i+15 in
  // lanes 0..15, read before this segment's carries of those rows are
  // written (a block writes only its own rows).
  auto load_res = [&](int i) -> uint32_t {
    return (i < Lpad) ? reinterpret_cast<const uint32_t *>(p + i)[lane & 3] : nb;
  };
  auto res_of = [](uint32_t v) -> uint4 {
    return make_uint4((uint32_t) __builtin_amdgcn_readlane((int) v, 0), (uint32_t) __builtin_amdgcn_readlane((int) v, 1),
                      (uint32_t) __builtin_amdgcn_readlane((int) v, 2), (uint32_t) __builtin_amdgcn_readlane((int) v, 3));
  };
  auto load_cin = [&](int i) -> uint32_t {
    return (!a.first && lane < 16 && i < Lpad) ? (uint32_t) cy[i + lane] : 0u;
  };
  // The incoming carries of rows i0 .. i0+15 (entering from rows i0-1 ..
  // i0+14) in lanes 0..15, high halves.
  auto carries_of = [&](uint32_t cin16) -> uint32_t {
    const uint32_t cin = cin16 << 16;
    const uint32_t x = __builtin_amdgcn_update_dpp(clast, cin, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
    clast = (uint32_t) __builtin_amdgcn_readlane((int) cin, 15);
    return x;
  };

  // A block's outgoing carries (lanes 48..63 of xo) are stored when the
  // next block starts, after its wait for the loads issued a block before,
  // so that wait does not include a store just issued.
  auto store_out = [&](uint32_t xo, int io) {
    if (!a.last && lane >= 48) cy[io + lane - 48] = (uint16_t) (xo >> 16);
  };

  uint4 c = res_of(load_res(0));
  uint32_t rv = load_res(16);
  uint32_t cv = load_cin(0);
  uint32_t xo = 0u;
  uint32_t e[QB * 4];
  load_row<Q>(tl, byte_of(c, 0), e);
  for (int i0 = 0; i0 < Lfull; i0 += 16) {
    const uint4 cn = res_of(rv);                      // block i0 + 16
    uint32_t x = carries_of(cv);
    if (i0 > 0) store_out(xo, i0 - 16);
    rv = load_res(i0 + 32);
    cv = load_cin(i0 + 16);
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      uint32_t en[QB * 4];
      load_row<Q>(tl, (rr < 15) ? byte_of(c, rr + 1) : byte_of(cn, 0), en);
      seg_row<Q, rr>(r, e, acc, x);
#pragma unroll
      for (int k = 0; k < Q; ++k) e[k] = en[k];
      row_fence<Q>(r);
      asm volatile("" : "+v"(acc), "+v"(x));
    });
    // the last row's outgoing carry into lane 63: lane 48 + k = row k
    xo = __builtin_amdgcn_update_dpp(r[Q - 1], x, 0x101, 0xf, 0xf, false);
    c = cn;
  }
  if (Lfull > 0) store_out(xo, Lfull - 16);
  const int rem = L - Lfull;                           // wave-uniform
  if (rem > 0) {
    uint32_t x = carries_of(cv);
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      if (rr < rem) {
        uint32_t en[QB * 4];
        if constexpr (rr < 15) load_row<Q>(tl, byte_of(c, rr + 1), en);
        seg_row<Q, rr>(r, e, acc, x);
        if constexpr (rr < 15) {
#pragma unroll
          for (int k = 0; k < Q; ++k) e[k] = en[k];
        }
        row_fence<Q>(r);
        asm volatile("" : "+v"(acc), "+v"(x));
      }
    });
    if constexpr ((Q & 1) != 0)
      if (rem & 1) acc = max2h(acc, r[0]);
    x = __builtin_amdgcn_update_dpp(r[Q - 1], x, 0x101, 0xf, 0xf, false);
    // lane 64 - rem + k = row k
    if (!a.last && lane >= 64 - rem) cy[Lfull + lane - (64 - rem)] = (uint16_t) (x >> 16);
  }

  const h2v ah = as_h2(acc);
  const float m = wave_max_f32(fmaxf((float) ah.x, (float) ah.y));
  if (lane == 0) {
    int Rp = (int) (m * 2048.0f + 0.5f);
    if (!a.first) Rp = max(Rp, (int) a.partial[s]);
    if (!a.last) {
      a.partial[s] = (uint16_t) Rp;
    } else {
      const uint32_t w = decide(Rp, sload(a.db.tjb + s), a.prof);
      a.prof.res[s] = (uint16_t) w;
      if ((w >> 8) == RES_SLOW)
        defer(w, a.prof_idx, s, a.slow, a.slow_count, a.slow_cap, a.jslow, a.jslow_count, a.jslow_cap);
    }
  }
}

template <int Q>
__global__ void __launch_bounds__(1024) msv_seg(SegArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int QB = (Q + 3) / 4;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int) (threadIdx.x >> 6));
  const int nwaves = blockDim.x >> 6;
  const int ntab = (a.db.R + 1) * QB * 64;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) lds[i] = a.tab[i];
  __syncthreads();
  const uint4 *tl = lds + lane;
  for (int s = blockIdx.x * nwaves + wave; s < a.db.n; s += gridDim.x * nwaves) seg_task<Q>(a, tl, s, lane);
}

// ---------------------------------------------------------------------------
// Stage 2a: J-aware pass for pairs whose J state may fire (msvfilter.c:
// 137-203 with exact integers).  Same f16 cell domain as the fast stage
// (W - B0 scaled by 2^-11, floor 0), plus per row: max with the current
// begin score B_i - B0, the wave's row maximum, the overflow test and the
// E -> J -> B update.  Cells below B0 are stored as B0; that changes neither
// max(M, xB) (xB >= B0) nor any row maximum that can move J past base, and
// the pairs queued here have final xJ > base, so the result is exact.
// ---------------------------------------------------------------------------

// One pair, one wave: tl = this lane's column of the 64-lane fast table
// (in LDS), p = the target's residue rows padded to 16 bytes with the null
// row.  Returns the result word (xJ, or RES_ERANGE << 8); xEmax receives the
// largest row maximum seen (stored cells, floor B0) up to the last row
// computed.
template <int Q>
__device__ __forceinline__ uint32_t jrow_task(const uint4 *tl, const uint8_t *p, int L, int tjb, const ProfDev &pd,
                                              int lane, int &xEmax) {
  constexpr int QB = (Q + 3) / 4;
  const int tjbm = (tjb + pd.tbm) & 255;                       // msvfilter.c:120
  const int B0 = max(0, pd.base - tjbm);                       // :124
  int xJ = 0, code = RES_OK;
  xEmax = B0;
  uint32_t bv = 0u;                                            // (xB - B0) / 2048 in both halves
  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;

  // Rows go in chunks of 16.  B changes only on a row whose xE - tec
  // exceeds max(base, xJ) (msvfilter.c:187-191), so each chunk is first
  // computed with B held at its value on entry, keeping only a per-lane
  // running maximum (no per-row wave reduction on the row chain).  If the
  // chunk's maximum xE stays within max(base, xJ) + tec, no row changed B
  // and every row was exact; the overflow test and xJ then follow from the
  // chunk maximum alone.  Otherwise the chunk is replayed row by row from
  // its entry state, and the next chunk starts row by row as well until a
  // chunk passes with B unchanged.  Targets are padded to 16 bytes with the
  // null row (all cells to the floor), so a partial last chunk is computed
  // whole in the first form: the extra rows leave the maximum unchanged.
  // The next chunk's residues are fetched while a chunk runs.
  const int Lpad = (L + 15) & ~15;
  const uint4 z4 = {0u, 0u, 0u, 0u};
  uint4 c = (L > 0) ? *reinterpret_cast<const uint4 *>(p) : z4;
  bool spec = true;
  for (int i0 = 0; i0 < L && code == RES_OK; i0 += 16) {
    const int i1 = i0 + 16;
    const uint4 cn = *reinterpret_cast<const uint4 *>(p + (i1 < Lpad ? i1 : 0));
    bool exact = !spec;
    if (spec) {
      uint32_t r0[Q];
#pragma unroll
      for (int j = 0; j < Q; ++j) r0[j] = r[j];
      uint32_t acc = 0u;
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) {
        uint32_t e[QB * 4];
        load_row<Q>(tl, byte_of(c, rr), e);
        const uint32_t sh = shift_slot<64>(r[Q - 1], lane);
#pragma unroll
        for (int j = Q - 1; j >= 1; --j) r[j] = cell(max2h(r[j - 1], bv), e[j]);   // :150-152
        r[0] = cell(max2h(sh, bv), e[0]);
        int j = 0;
#pragma unroll
        for (; j + 1 < Q; j += 2) acc = max3h(acc, r[j], r[j + 1]);                // :153
        if constexpr ((Q & 1) != 0) acc = max2h(acc, r[Q - 1]);
      }
      const h2v ah = as_h2(acc);
      const float m = wave_max_f32(fmaxf((float) ah.x, (float) ah.y));
      const int xEm = __builtin_amdgcn_readfirstlane((int) (m * 2048.0f + 0.5f)) + B0;
      xEmax = max(xEmax, xEm);
      if (xEm - pd.tec <= max(pd.base, xJ)) {
        if (xEm + pd.bias >= 255) { code = RES_ERANGE; break; }                    // :160-185
        xJ = max(xJ, max(0, xEm - pd.tec));                                        // :187-188
      } else {
#pragma unroll
        for (int j = 0; j < Q; ++j) r[j] = r0[j];
        exact = true;
      }
    }
    if (exact) {
      const int nr = min(16, L - i0);
      bool changed = false;
      uint32_t e[QB * 4];
      load_row<Q>(tl, byte_of(c, 0), e);
#pragma unroll
      for (int rr = 0; rr < 16; ++rr) {
        if (rr < nr && code == RES_OK) {
          uint32_t en[QB * 4];
          load_row<Q>(tl, (rr < 15) ? byte_of(c, rr + 1) : 0u, en);
          const uint32_t sh = shift_slot<64>(r[Q - 1], lane);
#pragma unroll
          for (int j = Q - 1; j >= 1; --j) r[j] = cell(max2h(r[j - 1], bv), e[j]);   // :150-152
          r[0] = cell(max2h(sh, bv), e[0]);
          uint32_t rm = 0u;
          int j = 0;
#pragma unroll
          for (; j + 1 < Q; j += 2) rm = max3h(rm, r[j], r[j + 1]);                  // :153
          if constexpr ((Q & 1) != 0) rm = max2h(rm, r[Q - 1]);
          const h2v rh = as_h2(rm);
          const float m = wave_max_f32(fmaxf((float) rh.x, (float) rh.y));
          const int xE = __builtin_amdgcn_readfirstlane((int) (m * 2048.0f + 0.5f)) + B0;
          xEmax = max(xEmax, xE);
#pragma unroll
          for (int k = 0; k < Q; ++k) e[k] = en[k];
          if (xE + pd.bias >= 255) { code = RES_ERANGE; break; }                     // :160-185
          xJ = max(xJ, max(0, xE - pd.tec));                                         // :187-188
          const int xB = max(0, max(pd.base, xJ) - tjbm);                            // :190-191
          const _Float16 bh = (_Float16) ((float) (xB - B0) * (1.0f / 2048.0f));
          const uint32_t b16 = __builtin_bit_cast(unsigned short, bh);
          const uint32_t nb = b16 | (b16 << 16);
          changed = changed || (nb != bv);
          bv = nb;
        }
      }
      spec = !changed;
    }
    c = cn;
  }
  return code == RES_ERANGE ? (RES_ERANGE << 8) : (uint32_t) xJ;
}

template <int Q>
__global__ void __launch_bounds__(256) msv_jrow(JArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int QB = (Q + 3) / 4;
  const int n = min(*a.count, a.cap);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  if ((int) blockIdx.x * nwaves >= n) return;      // no pair for this workgroup: leave before staging
  const int ntab = (a.db.R + 1) * QB * 64;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) lds[i] = a.tab[i];
  __syncthreads();
  for (int it = blockIdx.x * nwaves + wave; it < n; it += gridDim.x * nwaves) {
    const int64_t s = a.items[it].seq;
    int xEmax;
    const uint32_t w = jrow_task<Q>(lds + lane, a.db.res + a.db.off[s], a.db.len[s], a.db.tjb[s], a.prof, lane, xEmax);
    if (lane == 0) a.prof.res[s] = (uint16_t) w;
  }
}

// The J-aware pairs of many profiles in one launch (a batched search's J
// stage): one launch per profile lasts as long as its longest pair's row
// chain, so a profile's launch after another's on the same stream adds its
// whole chain; here every profile's pairs run side by side.  Workgroup b
// stages profile b % np's table; wave w of the profile's k-th workgroup
// takes pairs k * waves + w, then every wpp * waves further -- msv_jrow's
// static stride, with no device counter (a lane-0 atomic feeding a loop
// around this body was compiled into a nested loop that re-read a reset
// item index).  A workgroup whose profile has no pair for it leaves before
// staging.
template <int Q>
__device__ __forceinline__ void jrow_stride(const JManyArgs &a, const JProf &jp, const uint4 *lds, int n, int first,
                                            int stride, int lane) {
  for (int it = first; it < n; it += stride) {
    const int64_t s = jp.items[it].seq;
    int xEmax;
    const uint32_t w = jrow_task<Q>(lds + lane, a.db.res + a.db.off[s], a.db.len[s], a.db.tjb[s], jp.prof, lane, xEmax);
    if (lane == 0) jp.prof.res[s] = (uint16_t) w;
  }
}

__global__ void __launch_bounds__(512) msv_jrow_many(JManyArgs a) {
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int) (threadIdx.x >> 6));
  const int nwaves = blockDim.x >> 6;
  const JProf jp = a.profs[blockIdx.x % a.np];
  const int n = min(*jp.count, jp.cap);
  const int k = blockIdx.x / a.np;                       // the profile's k-th workgroup
  if (k * nwaves >= n) return;                           // no pair for it: leave before staging
  const int ntab = (a.db.R + 1) * ((jp.q + 3) / 4) * 64;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) lds[i] = jp.tab[i];
  __syncthreads();
  const int first = k * nwaves + wave, stride = a.wpp * nwaves;
  switch (jp.q) {
#define MSVG_JQ(q) case q: jrow_stride<q>(a, jp, lds, n, first, stride, lane); break;
    MSVG_JQ(1) MSVG_JQ(2) MSVG_JQ(3) MSVG_JQ(4) MSVG_JQ(5) MSVG_JQ(6) MSVG_JQ(7) MSVG_JQ(8)
    MSVG_JQ(9) MSVG_JQ(10) MSVG_JQ(11) MSVG_JQ(12) MSVG_JQ(13) MSVG_JQ(14) MSVG_JQ(15) MSVG_JQ(16)
    MSVG_JQ(17) MSVG_JQ(18) MSVG_JQ(19) MSVG_JQ(20) MSVG_JQ(21) MSVG_JQ(22) MSVG_JQ(23) MSVG_JQ(24)
    MSVG_JQ(25) MSVG_JQ(26) MSVG_JQ(27) MSVG_JQ(28) MSVG_JQ(29) MSVG_JQ(30) MSVG_JQ(31) MSVG_JQ(32)
#undef MSVG_JQ
    default: break;                    // the host sends q in 1..32 only
  }
}

// ---------------------------------------------------------------------------
// One-target composite for the per-call drop-in (msvfilter.c:78-213 for one
// dsq): the J-aware row loop above from the start of the target, then the
// composite's decision.  Before J first fires B is the constant B0, so up to
// that row the loop computes exactly the J-free cells of the fast stage:
// when J never fires (largest row maximum - tec <= base) the pair's word is
// decide() of that maximum, as in the fast stage; when it fires, the loop's
// own result is the J-aware stage's.  RES_SLOW words (the guard regime's
// sub-B0 maximum) are left to the caller.  One workgroup: all its threads
// stage the table in LDS, the first wave computes.
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(1024) msv_one(OneArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int QB = (Q + 3) / 4;
  const int ntab = (a.R + 1) * QB * 64;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) lds[i] = a.tab[i];
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  int xEmax;
  uint32_t w = jrow_task<Q>(lds + lane, a.res, a.L, a.tjb, a.prof, lane, xEmax);
  if ((w >> 8) != RES_ERANGE && xEmax - a.prof.tec <= a.prof.base) {
    const int tjbm = (a.tjb + a.prof.tbm) & 255;
    w = decide(xEmax - max(0, a.prof.base - tjbm), a.tjb, a.prof);
  }
  if (lane == 0) *a.out = w;
}

// ---------------------------------------------------------------------------
// Scan direction (hmmscan: many profiles against one query, or a batch of
// queries): the one-target composite above with one workgroup per profile.
// Its four waves stage the profile's table in LDS (four 16-byte reads in
// flight per lane), then take the queries in turn (wave w: queries w, w+4,
// ...); with one query three of them leave at once.  Workgroups are
// independent, so the hardware keeps as many resident per CU as their LDS
// allows (five for tables of 30 KiB) and starts the next as one ends.
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(256) msv_scan(ScanArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int QB = (Q + 3) / 4;
  const ScanProf *sp = a.profs + blockIdx.x;
  const uint4 *tab = sp->tab;
  // the table rows of the codes the queries hold (rowcode), then the null row
  constexpr int ROW = QB * 64;
  for (int r = 0; r <= a.R; ++r) {
    const uint4 *src = tab + (int) a.rowcode[r] * ROW;
    uint4 *dst = lds + r * ROW;
    for (int i = threadIdx.x; i < ROW; i += 256) dst[i] = src[i];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const ProfDev pd = sp->prof;
  const bool listed = a.wq_off != nullptr;
  const int k0 = listed ? a.wq_off[blockIdx.x] : 0;
  const int k1 = listed ? a.wq_off[blockIdx.x + 1] : a.nq;
  for (int k = k0 + wave; k < k1; k += 4) {       // waves without a query leave at once
    const int q = listed ? a.wq_idx[k] : k;
    const int tjb = a.qtjb[q];
    int xEmax;
    uint32_t w = jrow_task<Q>(lds + lane, a.res + a.qoff[q], a.qlen[q], tjb, pd, lane, xEmax);
    if ((w >> 8) != RES_ERANGE && xEmax - pd.tec <= pd.base) {
      const int tjbm = (tjb + pd.tbm) & 255;
      w = decide(xEmax - max(0, pd.base - tjbm), tjb, pd);
    }
    if (lane == 0) a.out[listed ? (int64_t) k : (int64_t) q * a.stride + blockIdx.x] = w;
  }
}

// ---------------------------------------------------------------------------
// Stage 2: exact byte MSV (msvfilter.c:112-210) for queued pairs.
// One wave per pair; the DP row lives in LDS, one dword column per lane.
// ---------------------------------------------------------------------------


#define SLOW_WAVES 4
#define SLOW_QMAX 128

__global__ void __launch_bounds__(64 * SLOW_WAVES) msv_slow(SlowArgs a) {
  __shared__ uint32_t dp[SLOW_WAVES][SLOW_QMAX][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int n = min(*a.count, a.cap);
  uint32_t *mydp = &dp[wave][0][0];

  for (int it = blockIdx.x * SLOW_WAVES + wave; it < n; it += gridDim.x * SLOW_WAVES) {
    const SlowItem item = a.items[it];
    const ProfDev pd = a.profs[item.prof];
    const int64_t s = item.seq;
    const int L = a.db.len[s];
    const uint8_t *p = a.db.res + a.db.off[s];
    const int Qs = pd.Qs;
    const int tjb = a.db.tjb[s];
    const int tjbm = (tjb + pd.tbm) & 255;                       // :120
    const u16x2 biasv = {(unsigned short) pd.bias, (unsigned short) pd.bias};
    const u16x2 c255 = {255, 255};

    for (int j = 0; j < Qs; ++j) mydp[j * 64 + lane] = 0u;       // :114
    int xJ = 0;                                                  // :115
    int xB = max(0, pd.base - tjbm);                             // :124
    int code = RES_OK;

    for (int i0 = 0; i0 < L; i0 += 16) {
      const uint4 c = *reinterpret_cast<const uint4 *>(p + i0);
      const int nr = min(16, L - i0);
      for (int rr = 0; rr < nr; ++rr) {
        const uint32_t x = byte_of(c, rr);
        const uint32_t *cp = pd.stab + (size_t) x * Qs * 64 + lane;
        const u16x2 xBv = {(unsigned short) xB, (unsigned short) xB};
        const uint32_t last = mydp[(Qs - 1) * 64 + lane];
        const uint32_t prev = __builtin_amdgcn_update_dpp(0u, last, 0x138, 0xf, 0xf, true);
        uint32_t mpv = __builtin_amdgcn_alignbit(last, prev, 16);  // :146, zero shifted in
        u16x2 xE = {0, 0};
        // Sixteen registers at a time: their costs and old DP values are read
        // together, so a row waits for one load latency per sixteen registers.
        for (int j0 = 0; j0 < Qs; j0 += 16) {
          uint32_t cst[16], old[16];
#pragma unroll
          for (int u = 0; u < 16; ++u)
            if (j0 + u < Qs) { cst[u] = cp[(j0 + u) * 64]; old[u] = mydp[(j0 + u) * 64 + lane]; }
#pragma unroll
          for (int u = 0; u < 16; ++u) {
            if (j0 + u < Qs) {
              u16x2 sv = __builtin_elementwise_max(as_u2(mpv), xBv);           // :150
              sv = __builtin_elementwise_min(sv + biasv, c255);                  // :151
              sv = __builtin_elementwise_sub_sat(sv, as_u2(cst[u]));             // :152
              xE = __builtin_elementwise_max(xE, sv);                            // :153
              mpv = old[u];
              mydp[(j0 + u) * 64 + lane] = u2_u32(sv);
            }
          }
        }
        const int m = wave_max_nonneg(max((int) xE.x, (int) xE.y));
        if (m + pd.bias >= 255) { code = RES_ERANGE; break; }                 // :160-185
        const int e = max(0, m - pd.tec);                                      // :187
        xJ = max(xJ, e);                                                       // :188
        xB = max(0, max(pd.base, xJ) - tjbm);                                  // :190-191
      }
      if (code == RES_ERANGE) break;
    }
    if (lane == 0) pd.res[s] = (uint16_t) (code == RES_ERANGE ? (RES_ERANGE << 8) : xJ);
  }
}

// ---------------------------------------------------------------------------
// Stage 3: usc, status and the F1 pass bit, in the caller's target order.
// ---------------------------------------------------------------------------


__global__ void __launch_bounds__(256) msv_finalize(FinArgs a) {
  const int p = blockIdx.y;
  const ProfDev pd = a.profs[p];
  const int64_t n = a.db.n;
  for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t) gridDim.x * blockDim.x) {
    const int64_t pos = a.sorted_pos[t];
    const uint32_t w = pd.res[pos];
    const uint32_t code = w >> 8;
    float usc;
    uint8_t st, ps;
    if (code == RES_ERANGE) {
      usc = __builtin_inff(); st = 16; ps = 1;               // eslERANGE, P = 0
    } else if (code == RES_OK) {
      const int xJ = (int) (w & 255u);
      usc = a.usc_tab[p * 511 + xJ - (int) a.db.tjb[pos] + 255];
      st = 0;
      ps = (xJ >= (int) a.thr[(int64_t) p * a.nL + a.lidx[pos]]) ? 1 : 0;
    } else {
      usc = __builtin_bit_cast(float, 0x7fc00000u); st = 255; ps = 0;  // unresolved: reported as an error
    }
    const int64_t o = (int64_t) p * n + t;
    if (a.usc) a.usc[o] = usc;
    if (a.status) a.status[o] = st;
    if (a.pass) a.pass[o] = ps;
  }
}

// ---------------------------------------------------------------------------
// Table builders and small utilities.
// ---------------------------------------------------------------------------
// Fast table: uint4 T[(R+1) * QB * 64]; see msv_device.h.
__global__ void build_fast_table(const uint8_t *rbv, int M, int Kp, const uint8_t *rowcode, int R,
                                 int G, int Q, int W, int bias, uint32_t *out) {
  const int QB = (Q + 3) / 4;
  const int64_t total = (int64_t) (R + 1) * QB * W * 4;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t) gridDim.x * blockDim.x) {
    const int e = (int) (i & 3);
    const int lane = (int) ((i >> 2) % W);
    const int64_t rest = (i >> 2) / W;
    const int jb = (int) (rest % QB);
    const int x = (int) (rest / QB);
    const int j = 4 * jb + e;
    const int t = lane % G;
    uint32_t word = 0;
    for (int h = 0; h < 2; ++h) {
      const int z = 2 * t + h;
      const int k = j + z * Q + 1;
      _Float16 v = (_Float16) -1.0f;
      if (x < R && j < Q && k <= M) {
        const int cost = rbv[(int64_t) rowcode[x] * M + (k - 1)];
        v = (_Float16) ((float) (bias - cost) * (1.0f / 2048.0f));
      }
      word |= (uint32_t) __builtin_bit_cast(unsigned short, v) << (16 * h);
    }
    out[i] = word;
  }
}

// Segment table: fast-table layout with G = 64 for positions k0+1 .. k0+128*Q.
__global__ void build_seg_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Q, int k0, int bias,
                                uint32_t *out) {
  const int QB = (Q + 3) / 4;
  const int64_t total = (int64_t) (R + 1) * QB * 64 * 4;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t) gridDim.x * blockDim.x) {
    const int e = (int) (i & 3);
    const int lane = (int) ((i >> 2) & 63);
    const int64_t rest = i >> 8;
    const int jb = (int) (rest % QB);
    const int x = (int) (rest / QB);
    const int j = 4 * jb + e;
    uint32_t word = 0;
    for (int h = 0; h < 2; ++h) {
      const int z = 2 * lane + h;
      const int k = k0 + j + z * Q + 1;
      _Float16 v = (_Float16) -1.0f;
      if (x < R && j < Q && k <= M) {
        const int cost = rbv[(int64_t) rowcode[x] * M + (k - 1)];
        v = (_Float16) ((float) (bias - cost) * (1.0f / 2048.0f));
      }
      word |= (uint32_t) __builtin_bit_cast(unsigned short, v) << (16 * h);
    }
    out[i] = word;
  }
}

// Slow table: uint32 C[(R+1) * Qs * 64]; dword = (cost of slot 2*lane, of slot 2*lane+1).
__global__ void build_slow_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Qs, uint32_t *out) {
  const int64_t total = (int64_t) (R + 1) * Qs * 64;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t) gridDim.x * blockDim.x) {
    const int lane = (int) (i & 63);
    const int64_t rest = i >> 6;
    const int j = (int) (rest % Qs);
    const int x = (int) (rest / Qs);
    uint32_t word = 0;
    for (int h = 0; h < 2; ++h) {
      const int z = 2 * lane + h;
      const int k = j + z * Qs + 1;
      uint32_t cost = MSVG_NULLCOST;
      if (x < R && k <= M) cost = rbv[(int64_t) rowcode[x] * M + (k - 1)];
      word |= cost << (16 * h);
    }
    out[i] = word;
  }
}

// Passing targets of one profile, compacted: index (caller order) and usc.
// The order of the output is unspecified; the host sorts it.
__global__ void __launch_bounds__(256) msv_compact(const uint8_t *pass, const float *usc, int64_t n,
                                                   int32_t *count, int32_t *idx, float *out_usc, int32_t cap) {
  for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t) gridDim.x * blockDim.x) {
    if (pass[t]) {
      const int k = atomicAdd(count, 1);
      if (k < cap) { idx[k] = (int32_t) t; out_usc[k] = usc[t]; }
    }
  }
}

// Zero n int32 words (used inside captured graphs in place of a memset node).
__global__ void fill_zero_i32(int32_t *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Host-side launch dispatch (called from msv_runtime.cpp).
// ---------------------------------------------------------------------------
typedef void (*fast_fn)(FastArgs);

template <int G, int Q>
static hipError_t launch_fast_t(const FastArgs &a, int grid, int block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((msv_fast<G, Q>), dim3(grid), dim3(block), lds, st, a);
  return hipGetLastError();
}

template <int G, int... Qs>
static hipError_t dispatch_q(int q, const FastArgs &a, int grid, int block, size_t lds, hipStream_t st,
                             std::integer_sequence<int, Qs...>) {
  hipError_t err = hipErrorInvalidValue;
  (void) ((q == Qs ? (err = launch_fast_t<G, Qs>(a, grid, block, lds, st), true) : false) || ...);
  return err;
}

// Q values compiled per lane-group size: every Q from 1 to 32.
#define MSVG_QSEQ std::integer_sequence<int, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, \
  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31, 32>

#define MSVG_SEGQ std::integer_sequence<int, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, \
  17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28>

template <int G, int Q>
static hipError_t set_lds_t() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_fast<G, Q>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}
template <int G, int... Qs>
static hipError_t set_lds_all(std::integer_sequence<int, Qs...>) {
  hipError_t err = hipSuccess;
  ((err = (err == hipSuccess ? set_lds_t<G, Qs>() : err)), ...);
  return err;
}

template <int Q>
static hipError_t launch_seg_t(const SegArgs &a, int grid, int block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((msv_seg<Q>), dim3(grid), dim3(block), lds, st, a);
  return hipGetLastError();
}
template <int... Qs>
static hipError_t dispatch_seg(int q, const SegArgs &a, int grid, int block, size_t lds, hipStream_t st,
                               std::integer_sequence<int, Qs...>) {
  hipError_t err = hipErrorInvalidValue;
  (void) ((q == Qs ? (err = launch_seg_t<Qs>(a, grid, block, lds, st), true) : false) || ...);
  return err;
}
template <int Q>
static hipError_t set_lds_seg_t() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_seg<Q>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}
template <int... Qs>
static hipError_t set_lds_seg(std::integer_sequence<int, Qs...>) {
  hipError_t err = hipSuccess;
  ((err = (err == hipSuccess ? set_lds_seg_t<Qs>() : err)), ...);
  return err;
}

extern "C" hipError_t msvk_launch_seg(int Q, const SegArgs *a, int grid, int block, size_t lds, hipStream_t st) {
  if (Q < 1 || Q > 28) return hipErrorInvalidValue;
  return dispatch_seg(Q, *a, grid, block, lds, st, MSVG_SEGQ{});
}

extern "C" hipError_t msvk_build_seg_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Q, int k0,
                                           int bias, uint32_t *out, hipStream_t st) {
  const int QB = (Q + 3) / 4;
  const int64_t total = (int64_t) (R + 1) * QB * 256;
  int grid = (int) ((total + 255) / 256);
  if (grid > 65535) grid = 65535;
  hipLaunchKernelGGL(build_seg_table, dim3(grid), dim3(256), 0, st, rbv, M, rowcode, R, Q, k0, bias, out);
  return hipGetLastError();
}

template <int Q>
static hipError_t launch_jrow_t(const JArgs &a, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((msv_jrow<Q>), dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}
template <int... Qs>
static hipError_t dispatch_jrow(int q, const JArgs &a, int grid, size_t lds, hipStream_t st,
                                std::integer_sequence<int, Qs...>) {
  hipError_t err = hipErrorInvalidValue;
  (void) ((q == Qs ? (err = launch_jrow_t<Qs>(a, grid, lds, st), true) : false) || ...);
  return err;
}
template <int Q>
static hipError_t set_lds_jrow_t() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_jrow<Q>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}
template <int... Qs>
static hipError_t set_lds_jrow(std::integer_sequence<int, Qs...>) {
  hipError_t err = hipSuccess;
  ((err = (err == hipSuccess ? set_lds_jrow_t<Qs>() : err)), ...);
  return err;
}

extern "C" hipError_t msvk_launch_jrow(int Q, const JArgs *a, int grid, size_t lds, hipStream_t st) {
  if (Q < 1 || Q > 32) return hipErrorInvalidValue;
  return dispatch_jrow(Q, *a, grid, lds, st, MSVG_QSEQ{});
}

// np * wpp workgroups of 512 threads; lds: the largest of the profiles' tables.
extern "C" hipError_t msvk_launch_jrow_many(const JManyArgs *a, size_t lds, hipStream_t st) {
  if (a->np < 1 || a->wpp < 1 || lds > 160 * 1024) return hipErrorInvalidValue;
  hipLaunchKernelGGL(msv_jrow_many, dim3(a->np * a->wpp), dim3(512), lds, st, *a);
  return hipGetLastError();
}

// Workgroups of msv_jrow_many resident per CU with this much LDS.
extern "C" hipError_t msvk_jrow_many_occupancy(size_t lds, int *per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, reinterpret_cast<const void *>(&msv_jrow_many), 512, lds);
}

template <int Q>
static hipError_t launch_one_t(const OneArgs &a, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((msv_one<Q>), dim3(1), dim3(1024), lds, st, a);
  return hipGetLastError();
}
template <int... Qs>
static hipError_t dispatch_one(int Q, const OneArgs &a, size_t lds, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t err = hipErrorInvalidValue;
  (void) ((Q == Qs ? (err = launch_one_t<Qs>(a, lds, st), true) : false) || ...);
  return err;
}
template <int Q>
static hipError_t set_lds_one_t() {
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_one<Q>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}
template <int... Qs>
static hipError_t set_lds_one(std::integer_sequence<int, Qs...>) {
  hipError_t err = hipSuccess;
  ((err = (err == hipSuccess ? set_lds_one_t<Qs>() : err)), ...);
  return err;
}

extern "C" hipError_t msvk_launch_one(int Q, const OneArgs *a, size_t lds, hipStream_t st) {
  if (Q < 1 || Q > 32) return hipErrorInvalidValue;
  return dispatch_one(Q, *a, lds, st, MSVG_QSEQ{});
}

// Register classes of the scan kernel: a profile's Q = ceil(M/128) is
// rounded up to the next class (the extra registers hold padding positions).
#define MSVG_SCANQ std::integer_sequence<int, 2, 4, 8, 12, 16, 20, 24, 28, 32>
template <int Q>
static hipError_t launch_scan_t(const ScanArgs &a, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((msv_scan<Q>), dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError();
}
template <int... Qs>
static hipError_t dispatch_scan(int Q, const ScanArgs &a, int grid, size_t lds, hipStream_t st,
                                std::integer_sequence<int, Qs...>) {
  hipError_t err = hipErrorInvalidValue;
  (void) ((Q == Qs ? (err = launch_scan_t<Qs>(a, grid, lds, st), true) : false) || ...);
  return err;
}
template <int... Qs>
static hipError_t set_lds_scan(std::integer_sequence<int, Qs...>) {
  hipError_t err = hipSuccess;
  ((err = (err == hipSuccess ? hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_scan<Qs>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)
                             : err)), ...);
  return err;
}

extern "C" hipError_t msvk_launch_scan(int Q, const ScanArgs *a, int grid, size_t lds, hipStream_t st) {
  if (grid < 1) return hipErrorInvalidValue;
  return dispatch_scan(Q, *a, grid, lds, st, MSVG_SCANQ{});
}

extern "C" hipError_t msvk_prepare() {
  hipError_t e;
  if ((e = set_lds_all<8>(MSVG_QSEQ{})) != hipSuccess) return e;
  if ((e = set_lds_seg(MSVG_SEGQ{})) != hipSuccess) return e;
  if ((e = set_lds_jrow(MSVG_QSEQ{})) != hipSuccess) return e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_jrow_many),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess) return e;
  if ((e = set_lds_one(MSVG_QSEQ{})) != hipSuccess) return e;
  if ((e = set_lds_scan(MSVG_SCANQ{})) != hipSuccess) return e;
  if ((e = set_lds_all<16>(MSVG_QSEQ{})) != hipSuccess) return e;
  if ((e = set_lds_all<32>(MSVG_QSEQ{})) != hipSuccess) return e;
  return set_lds_all<64>(MSVG_QSEQ{});
}

extern "C" hipError_t msvk_launch_fast(int G, int Q, const FastArgs *a, int grid, int block, size_t lds,
                                       hipStream_t st) {
  if (Q < 1 || Q > 32) return hipErrorInvalidValue;
  switch (G) {
    case 8: return dispatch_q<8>(Q, *a, grid, block, lds, st, MSVG_QSEQ{});
    case 16: return dispatch_q<16>(Q, *a, grid, block, lds, st, MSVG_QSEQ{});
    case 32: return dispatch_q<32>(Q, *a, grid, block, lds, st, MSVG_QSEQ{});
    case 64: return dispatch_q<64>(Q, *a, grid, block, lds, st, MSVG_QSEQ{});
    default: return hipErrorInvalidValue;
  }
}

extern "C" hipError_t msvk_launch_slow(const SlowArgs *a, int grid, hipStream_t st) {
  hipLaunchKernelGGL(msv_slow, dim3(grid), dim3(64 * SLOW_WAVES), 0, st, *a);
  return hipGetLastError();
}

extern "C" hipError_t msvk_launch_finalize(const FinArgs *a, int np, int64_t n, hipStream_t st) {
  int gx = (int) ((n + 255) / 256);
  if (gx > 4096) gx = 4096;
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(msv_finalize, dim3(gx, np), dim3(256), 0, st, *a);
  return hipGetLastError();
}

extern "C" hipError_t msvk_build_fast_table(const uint8_t *rbv, int M, int Kp, const uint8_t *rowcode, int R, int G,
                                            int Q, int W, int bias, uint32_t *out, hipStream_t st) {
  if (W < G || W > 64 || W % G != 0) return hipErrorInvalidValue;
  const int QB = (Q + 3) / 4;
  const int64_t total = (int64_t) (R + 1) * QB * W * 4;
  int grid = (int) ((total + 255) / 256);
  if (grid > 65535) grid = 65535;
  hipLaunchKernelGGL(build_fast_table, dim3(grid), dim3(256), 0, st, rbv, M, Kp, rowcode, R, G, Q, W, bias, out);
  return hipGetLastError();
}

extern "C" hipError_t msvk_build_slow_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int Qs,
                                            uint32_t *out, hipStream_t st) {
  const int64_t total = (int64_t) (R + 1) * Qs * 64;
  int grid = (int) ((total + 255) / 256);
  if (grid > 65535) grid = 65535;
  hipLaunchKernelGGL(build_slow_table, dim3(grid), dim3(256), 0, st, rbv, M, rowcode, R, Qs, out);
  return hipGetLastError();
}

extern "C" hipError_t msvk_compact(const uint8_t *pass, const float *usc, int64_t n, int32_t *count, int32_t *idx,
                                   float *out_usc, int32_t cap, hipStream_t st) {
  int grid = (int) ((n + 255) / 256);
  if (grid > 8192) grid = 8192;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL(msv_compact, dim3(grid), dim3(256), 0, st, pass, usc, n, count, idx, out_usc, cap);
  return hipGetLastError();
}

extern "C" hipError_t msvk_fill_zero(int32_t *p, int n, hipStream_t st) {
  hipLaunchKernelGGL(fill_zero_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

// LDS read-bandwidth probe (measurement only): 16 waves per CU stream
// conflict-free ds_read_b128 over a 64 KiB table, 8 per iteration; the
// bytes read are grid * 1024 * iters * 8 * 16.  Every index stays inside
// the table; out receives one word per thread so the reads are kept.
__global__ void __launch_bounds__(1024) lds_probe(uint4 *out, int iters) {
  __shared__ uint4 t[4096];
  for (int i = threadIdx.x; i < 4096; i += blockDim.x) t[i] = make_uint4(i, i + 1, i + 2, i + 3);
  __syncthreads();
  uint4 acc = make_uint4(0u, 0u, 0u, 0u);
  int idx = threadIdx.x;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const uint4 v = t[(idx + u * 512) & 4095];
      acc.x ^= v.x; acc.y ^= v.y; acc.z ^= v.z; acc.w ^= v.w;
    }
    idx = (idx + 64) & 4095;
  }
  out[(size_t) blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

extern "C" hipError_t msvk_lds_probe(void *out, int grid, int iters, hipStream_t st) {
  hipLaunchKernelGGL(lds_probe, dim3(grid), dim3(1024), 0, st, static_cast<uint4 *>(out), iters);
  return hipGetLastError();
}

// Packed-form probe (measurement only): the J-free recurrence of the
// batched kernel's 64-lane layout (one target per wave, Q = 16 registers =
// 2048 model positions, emissions from an LDS table read one row ahead,
// the one-slot shift) over L rows, in two encodings of the same cells:
//   FORM 0, the kernels' form: W - B0 scaled by 2^-11 in f16;
//     v_pk_add_f16 ... clamp per register, v_pk_maximum3_f16 per two;
//   FORM 1, unsigned 16-bit units: v_pk_add_u16 of the bias, v_pk_sub_u16
//     ... clamp of the cost (the floor at B0), v_pk_max_u16 per register.
// Table entry (x, block, lane) as the fast-stage table (21 rows, 4 blocks):
// FORM 0 holds (bias - cost) * 2^-11, FORM 1 the costs.  Residues: L bytes
// per wave (L a multiple of 16), codes below 21.  out[wave] = the maximum
// cell in units; both forms agree while it stays below 2048.
template <int FORM>
__global__ void __launch_bounds__(1024) forms_probe(const uint8_t *res, int L, const uint4 *tab, int tab_u4,
                                                    uint32_t bias2, uint32_t *out) {
  extern __shared__ uint4 lds[];
  constexpr int Q = 16, QB = 4;
  for (int i = threadIdx.x; i < tab_u4; i += blockDim.x) lds[i] = tab[i];
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = (int) blockIdx.x * (int) (blockDim.x / 64) + (int) (threadIdx.x / 64);
  const uint8_t *p = res + (size_t) wave * L;
  const uint4 *tl = lds + lane;
  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;
  uint32_t acc = 0u;
  uint4 c = *reinterpret_cast<const uint4 *>(p);
  uint32_t e[QB * 4];
  load_row<Q>(tl, (uint32_t) __builtin_amdgcn_readfirstlane((int) byte_of(c, 0)), e);
  for (int i0 = 0; i0 < L; i0 += 16) {
    const int i1 = (i0 + 16 < L) ? i0 + 16 : i0;
    const uint4 cn = *reinterpret_cast<const uint4 *>(p + i1);
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      uint32_t en[QB * 4];
      const uint32_t xn = (rr < 15) ? byte_of(c, rr + 1) : byte_of(cn, 0);
      load_row<Q>(tl, (uint32_t) __builtin_amdgcn_readfirstlane((int) xn), en);
      const uint32_t in0 = shift_slot<64>(r[Q - 1], lane);
      if constexpr (FORM == 0) {
#pragma unroll
        for (int j = Q - 1; j > 0; --j) r[j] = cell(r[j - 1], e[j]);
        r[0] = cell(in0, e[0]);
#pragma unroll
        for (int j = 0; j < Q; j += 2) acc = max3h(acc, r[j], r[j + 1]);
      } else {
        const u16x2 b = as_u2(bias2);
#pragma unroll
        for (int j = Q - 1; j > 0; --j)
          r[j] = u2_u32(__builtin_elementwise_sub_sat(as_u2(r[j - 1]) + b, as_u2(e[j])));
        r[0] = u2_u32(__builtin_elementwise_sub_sat(as_u2(in0) + b, as_u2(e[0])));
#pragma unroll
        for (int j = 0; j < Q; ++j) acc = u2_u32(__builtin_elementwise_max(as_u2(acc), as_u2(r[j])));
      }
#pragma unroll
      for (int k = 0; k < Q; ++k) e[k] = en[k];
      row_fence<Q>(r);
      asm volatile("" : "+v"(acc));
    });
    c = cn;
  }
  int m;
  if constexpr (FORM == 0) {
    const h2v a = as_h2(acc);
    m = (int) (wave_max_f32(fmaxf((float) a.x, (float) a.y)) * 2048.0f + 0.5f);
  } else {
    const u16x2 a = as_u2(acc);
    m = (int) (wave_max_f32((float) (a.x > a.y ? a.x : a.y)) + 0.5f);
  }
  if (lane == 0) out[wave] = (uint32_t) m;
}

extern "C" hipError_t msvk_forms_probe(int form, const uint8_t *res, int L, const void *tab, int tab_u4,
                                       uint32_t bias2, uint32_t *out, int grid, hipStream_t st) {
  const size_t lds = (size_t) tab_u4 * 16;
  hipError_t e;
  if (form == 0) {
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&forms_probe<0>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(forms_probe<0>, dim3(grid), dim3(1024), lds, st, res, L, static_cast<const uint4 *>(tab),
                       tab_u4, bias2, out);
  } else {
    if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&forms_probe<1>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds)) != hipSuccess) return e;
    hipLaunchKernelGGL(forms_probe<1>, dim3(grid), dim3(1024), lds, st, res, L, static_cast<const uint4 *>(tab),
                       tab_u4, bias2, out);
  }
  return hipGetLastError();
}
