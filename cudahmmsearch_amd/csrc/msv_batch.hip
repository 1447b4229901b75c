// msv_batch.hip -- the batched J-free pass: many profiles against one resident
// database in one persistent launch per lane-group size G.
//
// The recurrence and decision are those of msv_fast (msv_kernels.hip): the
// J-free maximum over all cells with the begin score B0 as a floor, in packed
// f16 (two cells per VGPR, the clamp bit is the floor), then SSV's decision
// rules (impl_sse/ssvfilter.c:1370-1410; decide() in msv_cells.h).  What
// differs is the work per residue row and the schedule:
//
//  * Row addresses cost one VALU instruction.  G = 64: a wave scores one
//    target, so its residues are wave-uniform; they come through the scalar
//    cache and the table offset of a row is scalar arithmetic plus one
//    v_add.  G = 8, 16, 32: the table is laid out [register block][32 rows]
//    [TW lanes] (msv_device.h), so one v_perm_b32 turns a residue byte into
//    the row's byte address and the register block is the read's immediate
//    offset.
//  * For an odd register count the leftover register of a row is paired with
//    the same register of the next row in one v_pk_maximum3_f16.
//  * Each profile has a queue of target chunks (the database's guided chunk
//    list, longest targets first).  Workgroups start spread over the profiles
//    in proportion to their model length, drain their profile's queue wave
//    by wave, and move to the next profile with chunks left when it is
//    empty, restaging the LDS table; so every profile's longest targets start
//    at once and one profile's tail overlaps the others' work.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "msv_device.h"
#include "msv_cells.h"

namespace {

// LDS byte address -> pointer.  msv_batch declares no static LDS, so its
// dynamic LDS (the table) starts at address 0 and a table byte offset is
// its LDS address; building the pointer from the integer keeps the register
// block's offset in the read's immediate field (an extern __shared__ base
// is added per read otherwise).
typedef __attribute__((address_space(3))) const char lds_byte;
typedef uint32_t lds_v4 __attribute__((ext_vector_type(4)));
typedef uint32_t lds_v2 __attribute__((ext_vector_type(2)));

// Emissions of one residue row for the Q registers of this lane: QB entries
// of 16 bytes, STRIDE bytes apart from LDS byte address addr; the last entry
// is read only as wide as it is used.
template <int Q, uint32_t STRIDE>
__device__ __forceinline__ void load_at(uint32_t addr, uint32_t (&e)[((Q + 3) / 4) * 4]) {
  constexpr int QB = (Q + 3) / 4, TAIL = Q - 4 * (QB - 1);
  lds_byte *base = (lds_byte *) (uintptr_t) addr;
#pragma unroll
  for (int jb = 0; jb < QB - 1; ++jb) {
    const lds_v4 v = *reinterpret_cast<__attribute__((address_space(3))) const lds_v4 *>(base + jb * STRIDE);
    e[4 * jb + 0] = v.x; e[4 * jb + 1] = v.y; e[4 * jb + 2] = v.z; e[4 * jb + 3] = v.w;
  }
  lds_byte *lp = base + (QB - 1) * STRIDE;
  constexpr int b = 4 * (QB - 1);
  if constexpr (TAIL == 4) {
    const lds_v4 v = *reinterpret_cast<__attribute__((address_space(3))) const lds_v4 *>(lp);
    e[b] = v.x; e[b + 1] = v.y; e[b + 2] = v.z; e[b + 3] = v.w;
  } else if constexpr (TAIL == 3) {
    const lds_v2 v = *reinterpret_cast<__attribute__((address_space(3))) const lds_v2 *>(lp);
    e[b] = v.x; e[b + 1] = v.y;
    e[b + 2] = reinterpret_cast<__attribute__((address_space(3))) const uint32_t *>(lp)[2];
  } else if constexpr (TAIL == 2) {
    const lds_v2 v = *reinterpret_cast<__attribute__((address_space(3))) const lds_v2 *>(lp);
    e[b] = v.x; e[b + 1] = v.y;
  } else {
    e[b] = *reinterpret_cast<__attribute__((address_space(3))) const uint32_t *>(lp);
  }
}

// One DP row of the J-free recurrence on the Q registers of this lane (in
// place, descending, so register j reads register j-1's value of the row
// before), and the row's cells into the running maximum, two registers per
// v_pk_maximum3_f16.  Row RR of a block of 16, for odd Q: an even row leaves
// out register 0, and the odd row after it maxes that value (the register
// keeps it until the odd row's last update) with its own register Q-1 (for
// Q = 1 that is register 0 after its update).
template <int G, int Q, int RR>
__device__ __forceinline__ void row_step(uint32_t (&r)[Q], const uint32_t *e, uint32_t &acc, int lane) {
  const uint32_t sh = shift_slot<G>(r[Q - 1], lane);
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

// Where a pair's result word goes.  A search: the profile's result array,
// and a pair whose J state may fire or that needs the byte-exact pass is
// queued.  A grouped scan: the word only, in the profile's column of the
// [target][profile] word array (the host lists such pairs).
struct BatOut {
  uint16_t *res;
  SlowItem *slow;
  int32_t *slow_count;
  SlowItem *jslow;
  int32_t *jslow_count;
  int32_t prof_idx, slow_cap, jslow_cap;
  __device__ __forceinline__ void operator()(int64_t s, uint32_t w) const {
    res[s] = (uint16_t) w;
    if ((w >> 8) == RES_SLOW) defer(w, prof_idx, s, slow, slow_count, slow_cap, jslow, jslow_count, jslow_cap);
  }
};

struct ScanOut {
  uint16_t *col;
  int64_t stride;
  __device__ __forceinline__ void operator()(int64_t s, uint32_t w) const { col[s * stride] = (uint16_t) w; }
};

// G = 64: one target per wave (s is wave-uniform).
template <int Q, class Out>
__device__ __forceinline__ void task64(const DbDev &db, const ProfDev &pd, const uint4 *lds, int s, int lane,
                                       const Out &out) {
  constexpr int QB = (Q + 3) / 4;
  const int L = sload(db.len + s);
  const int Lpad = (L + 15) & ~15;
  const uint8_t *p = db.res + sload(db.off + s);
  const uint4 *tl = lds + lane;
  const uint32_t nb = (uint32_t) db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);

  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;
  uint32_t acc = 0u;

  // Residues 16 at a time through the scalar cache, the next block fetched
  // while the current one runs; each row's emissions are read one row ahead.
  // Full blocks of 16 rows first, then the L % 16 rows of the last block
  // (the padding rows after them would only add cells at the floor).  (A
  // block past the target's end reads its first bytes again: in bounds.)
  const int Lfull = L & ~15;
  uint4 c = (Lpad > 0) ? sload(reinterpret_cast<const uint4 *>(p)) : null4;
  uint32_t e[QB * 4];
  load_row<Q>(tl, byte_of(c, 0), e);
  for (int i0 = 0; i0 < Lfull; i0 += 16) {
    const int i1 = i0 + 16;
    const int inext = __builtin_amdgcn_readfirstlane(i1 < Lpad ? i1 : 0);
    const uint4 cl = sload(reinterpret_cast<const uint4 *>(p + inext));
    const uint4 cn = (i1 < Lpad) ? cl : null4;
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      uint32_t en[QB * 4];
      load_row<Q>(tl, (rr < 15) ? byte_of(c, rr + 1) : byte_of(cn, 0), en);
      row_step<64, Q, rr>(r, e, acc, lane);
#pragma unroll
      for (int k = 0; k < Q; ++k) e[k] = en[k];
      row_fence<Q>(r);
      asm volatile("" : "+v"(acc));
    });
    c = cn;
  }
  const int rem = L - Lfull;                           // wave-uniform
  if (rem > 0) {
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      if (rr < rem) {
        uint32_t en[QB * 4];
        if constexpr (rr < 15) load_row<Q>(tl, byte_of(c, rr + 1), en);
        row_step<64, Q, rr>(r, e, acc, lane);
        if constexpr (rr < 15) {
#pragma unroll
          for (int k = 0; k < Q; ++k) e[k] = en[k];
        }
        row_fence<Q>(r);
        asm volatile("" : "+v"(acc));
      }
    });
    // odd Q: a last row of even index left its register 0 for a next row
    if constexpr ((Q & 1) != 0)
      if (rem & 1) acc = max2h(acc, r[0]);
  }

  const h2v ah = as_h2(acc);
  const float m = wave_max_f32(fmaxf((float) ah.x, (float) ah.y));
  if (lane == 0) {
    const int Rp = (int) (m * 2048.0f + 0.5f);
    out(s, decide(Rp, sload(db.tjb + s), pd));
  }
}

// The 16 residue bytes of c as row-address bytes: x for TW = 16 and
// 2x + (column >= 16) for TW = 32 (codes are below 32, so no byte carries).
template <int TW>
__device__ __forceinline__ uint4 addr_bytes(const uint4 &c, uint32_t half) {
  if constexpr (TW == 32)
    return make_uint4((c.x << 1) | half, (c.y << 1) | half, (c.z << 1) | half, (c.w << 1) | half);
  else
    return c;
}

// Byte address of row rr of a 16-row block: byte 1 = the row byte, byte 0 =
// the lane's column offset (v_perm_b32).
template <int RR>
__device__ __forceinline__ uint32_t row_addr(const uint4 &w, uint32_t col) {
  constexpr int k = RR >> 2;
  const uint32_t word = (k == 0) ? w.x : (k == 1) ? w.y : (k == 2) ? w.z : w.w;
  return __builtin_amdgcn_perm(word, col, 0x0c0c0000u | ((4u + (uint32_t) (RR & 3)) << 8));
}

// G = 8, 16, 32: 64/G targets per wave, s is this lane group's target.
template <int G, int Q, class Out>
__device__ __forceinline__ void task_narrow(const DbDev &db, const ProfDev &pd, const uint4 *lds, int64_t s,
                                            int lane, const Out &out) {
  constexpr int QB = (Q + 3) / 4;
  constexpr int TW = G < 16 ? 16 : G;
  constexpr uint32_t JB = 32u * TW * 16u;               // bytes per register block
  const bool valid = s < db.n;
  const int L = valid ? db.len[s] : 0;
  const int Lpad = (L + 15) & ~15;
  const uint8_t *p = db.res + (valid ? db.off[s] : 0);

  int Lmax = L;                                         // the wave's longest target
#pragma unroll
  for (int o = 32; o >= G; o >>= 1) Lmax = max(Lmax, __shfl_xor(Lmax, o, 64));
  Lmax = __builtin_amdgcn_readfirstlane(Lmax);
  const int Lfull = Lmax & ~15;

  const uint32_t nb = (uint32_t) db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);
  const uint32_t col = (uint32_t) (lane & 15) * 16u;
  const uint32_t half = (TW == 32) ? (uint32_t) ((lane >> 4) & 1) * 0x01010101u : 0u;

  uint32_t r[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) r[j] = 0u;
  uint32_t acc = 0u;

  uint4 cnext = *reinterpret_cast<const uint4 *>(p);
  uint4 cw = addr_bytes<TW>((0 < Lpad) ? cnext : null4, half);
  uint32_t e[QB * 4];
  load_at<Q, JB>(row_addr<0>(cw, col), e);
  for (int i0 = 0; i0 < Lfull; i0 += 16) {
    const int i1 = i0 + 16;
    cnext = *reinterpret_cast<const uint4 *>(p + (i1 < Lpad ? i1 : 0));
    const uint4 cnw = addr_bytes<TW>((i1 < Lpad) ? cnext : null4, half);
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      uint32_t en[QB * 4];
      if constexpr (rr < 15) load_at<Q, JB>(row_addr<rr + 1>(cw, col), en);
      else load_at<Q, JB>(row_addr<0>(cnw, col), en);
      row_step<G, Q, rr>(r, e, acc, lane);
#pragma unroll
      for (int k = 0; k < Q; ++k) e[k] = en[k];
      row_fence<Q>(r);
      asm volatile("" : "+v"(acc));
    });
    cw = cnw;
  }
  const int rem = Lmax - Lfull;                        // wave-uniform; shorter targets read null rows
  if (rem > 0) {
    static_for<0, 16>([&](auto rrc) {
      constexpr int rr = decltype(rrc)::value;
      if (rr < rem) {
        uint32_t en[QB * 4];
        if constexpr (rr < 15) load_at<Q, JB>(row_addr<rr + 1>(cw, col), en);
        row_step<G, Q, rr>(r, e, acc, lane);
        if constexpr (rr < 15) {
#pragma unroll
          for (int k = 0; k < Q; ++k) e[k] = en[k];
        }
        row_fence<Q>(r);
        asm volatile("" : "+v"(acc));
      }
    });
    if constexpr ((Q & 1) != 0)
      if (rem & 1) acc = max2h(acc, r[0]);
  }

  const h2v ah = as_h2(acc);
  float m = fmaxf((float) ah.x, (float) ah.y);
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (valid && (lane % G) == 0) {
    const int Rp = (int) (m * 2048.0f + 0.5f);
    out(s, decide(Rp, db.tjb[s], pd));
  }
}

// Drain one profile's chunk queue with this wave.  Every lane of the wave is
// active here; lane 0 takes the chunk and readlane(., 0) hands it to all.
template <int G, int Q>
__device__ __forceinline__ void drain(const BatArgs &a, const BatProf &bp, const uint4 *lds, int lane) {
  constexpr int NG = 64 / G;
  const int g = lane / G;
  const BatOut out{bp.pd.res, a.slow, a.slow_count, bp.jslow, bp.jslow_count, bp.prof_idx, a.slow_cap, a.jslow_cap};
  for (;;) {
    int c = 0;
    if (lane == 0) c = atomicAdd(bp.counter, 1);
    c = __builtin_amdgcn_readlane(c, 0);
    if (c >= a.nchunks) break;
    const int t0 = sload(a.chunk_start + c), t1 = sload(a.chunk_start + c + 1);
    for (int t = t0; t < t1; ++t) {
      if constexpr (G == 64) task64<Q>(a.db, bp.pd, lds, t, lane, out);
      else task_narrow<G, Q>(a.db, bp.pd, lds, (int64_t) t * NG + g, lane, out);
    }
  }
}

// Register counts built per lane-group size: the narrow layout's immediate
// offsets pass 16 bits beyond 4 register blocks for TW = 32, and past 16 (G =
// 8) or 24 (G = 16) registers the narrow tasks run out of VGPRs at 1024
// threads per workgroup.  The host moves such profiles to a wider G.
template <int G>
__host__ __device__ constexpr int bat_qmax() { return G == 64 ? 32 : G == 16 ? 24 : 16; }

template <int G, int Q>
__device__ __forceinline__ void drain_if(const BatArgs &a, const BatProf &bp, const uint4 *lds, int lane) {
  if constexpr (Q <= bat_qmax<G>()) drain<G, Q>(a, bp, lds, lane);
}

}  // namespace

template <int G>
__global__ void __launch_bounds__(1024) msv_batch(BatArgs a) {
  extern __shared__ uint4 lds[];
  const int lane = threadIdx.x & 63;
  int32_t *slot = reinterpret_cast<int32_t *>(lds + a.slot_u4);
  int p = sload(a.first + blockIdx.x);
  // Every thread of the workgroup runs the same sequence of barriers: p is
  // uniform (written by thread 0, read by all after a barrier), and each
  // wave's drain() ends once its profile's queue is empty.
  while (p >= 0) {
    const BatProf bp = sload(a.profs + p);
    for (int i = threadIdx.x; i < bp.tab_u4; i += blockDim.x) lds[i] = bp.tab[i];
    __syncthreads();
    switch (bp.pd.Q) {
#define MSVG_BQ(q) case q: drain_if<G, q>(a, bp, lds, lane); break;
      MSVG_BQ(1) MSVG_BQ(2) MSVG_BQ(3) MSVG_BQ(4) MSVG_BQ(5) MSVG_BQ(6) MSVG_BQ(7) MSVG_BQ(8)
      MSVG_BQ(9) MSVG_BQ(10) MSVG_BQ(11) MSVG_BQ(12) MSVG_BQ(13) MSVG_BQ(14) MSVG_BQ(15) MSVG_BQ(16)
      MSVG_BQ(17) MSVG_BQ(18) MSVG_BQ(19) MSVG_BQ(20) MSVG_BQ(21) MSVG_BQ(22) MSVG_BQ(23) MSVG_BQ(24)
      MSVG_BQ(25) MSVG_BQ(26) MSVG_BQ(27) MSVG_BQ(28) MSVG_BQ(29) MSVG_BQ(30) MSVG_BQ(31) MSVG_BQ(32)
#undef MSVG_BQ
      default:
        // Not built for this G (the host never sends such a profile): empty
        // the queue without scoring so no workgroup keeps coming back to it.
        for (;;) {
          int c = 0;
          if (lane == 0) c = atomicAdd(bp.counter, 1);
          if (__builtin_amdgcn_readlane(c, 0) >= a.nchunks) break;
        }
        break;
    }
    __syncthreads();                   // no wave reads the table any more
    if (threadIdx.x == 0) {
      // The next profile (in order after p) whose queue still has chunks.
      // The counters are read with device-scope atomic loads; a queue that
      // empties after it is picked costs one restage and one pass here.
      int nx = -1;
      for (int k = 1; k <= a.np && nx < 0; ++k) {
        const int q = (p + k) % a.np;
        const int32_t *cq = a.profs[q].counter;
        if (__hip_atomic_load(cq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < a.nchunks) nx = q;
      }
      *slot = nx;
    }
    __syncthreads();
    p = *slot;
  }
}

// ---------------------------------------------------------------------------
// Grouped scan (the hmmscan direction, a batch of queries): workgroup b
// scores profile b % np against the tasks of its wave slots, with the row
// loops above.  Not persistent and no counters: the host's slot schedule
// (the same for every profile of the launch) says which tasks each wave
// takes.  QB is the launch's register-block count; pd.Q picks among the four
// register counts it holds.
// ---------------------------------------------------------------------------
namespace {

template <int G, int Q>
__device__ __forceinline__ void scan_slot(const ScanGroupArgs &a, const ProfDev &pd, const uint4 *lds, int slot,
                                          int lane, const ScanOut &out) {
  if constexpr (Q <= bat_qmax<G>()) {
    constexpr int NG = 64 / G;
    const int i1 = sload(a.soff + slot + 1);
    for (int i = sload(a.soff + slot); i < i1; ++i) {
      const int t = sload(a.stask + i);
      if constexpr (G == 64) task64<Q>(a.db, pd, lds, t, lane, out);
      else task_narrow<G, Q>(a.db, pd, lds, (int64_t) t * NG + lane / G, lane, out);
    }
  }
}

}  // namespace

template <int G, int QB>
__global__ void __launch_bounds__(1024) msv_scan_group(ScanGroupArgs a) {
  extern __shared__ uint4 lds[];
  constexpr int TW = G < 16 ? 16 : G;
  if (a.prio) __builtin_amdgcn_s_setprio(2);
  const int lane = threadIdx.x & 63;
  const int sp = (int) (blockIdx.x % (unsigned) a.np);
  const int split = (int) (blockIdx.x / (unsigned) a.np);
  const ProfDev pd = sload(a.profs + sp);
  const int R1 = a.db.R + 1;
  if constexpr (G == 64) {                 // [row][jb][64]: whole rows
    constexpr int row = QB * 64;
    for (int i = threadIdx.x; i < R1 * row; i += blockDim.x) {
      const int r = i / row;
      lds[i] = pd.ftab[(int) a.rowcode[r] * row + (i - r * row)];
    }
  } else {                                 // [jb][32 rows][TW]: R + 1 rows of each block
    const int blk = R1 * TW;
    for (int i = threadIdx.x; i < QB * blk; i += blockDim.x) {
      const int jb = i / blk, rem = i - jb * blk;
      const int r = rem / TW, col = rem - r * TW;
      lds[(jb * 32 + r) * TW + col] = pd.ftab[(jb * 32 + (int) a.rowcode[r]) * TW + col];
    }
  }
  __syncthreads();
  const int slot = __builtin_amdgcn_readfirstlane(split * (int) (blockDim.x >> 6) + (int) (threadIdx.x >> 6));
  const ScanOut out{a.out + sp, a.stride};
  switch (pd.Q - 4 * (QB - 1)) {
    case 1: scan_slot<G, 4 * QB - 3>(a, pd, lds, slot, lane, out); break;
    case 2: scan_slot<G, 4 * QB - 2>(a, pd, lds, slot, lane, out); break;
    case 3: scan_slot<G, 4 * QB - 1>(a, pd, lds, slot, lane, out); break;
    case 4: scan_slot<G, 4 * QB>(a, pd, lds, slot, lane, out); break;
    default: break;                        // not a register count of this launch: the host never sends one
  }
}

// Batched-pass table for G = 8, 16, 32 (layout in msv_device.h).
__global__ void build_bat_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int G, int Q, int TW,
                                int bias, uint32_t *out) {
  const int QB = (Q + 3) / 4;
  const int64_t total = (int64_t) QB * 32 * TW * 4;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t) gridDim.x * blockDim.x) {
    const int e = (int) (i & 3);
    const int column = (int) ((i >> 2) % TW);
    const int64_t rest = (i >> 2) / TW;
    const int x = (int) (rest % 32);
    const int jb = (int) (rest / 32);
    const int j = 4 * jb + e;
    const int t = column % G;
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

extern "C" hipError_t msvk_prepare_batch() {
  hipError_t e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_batch<8>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess) return e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_batch<16>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess) return e;
  if ((e = hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_batch<32>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024)) != hipSuccess) return e;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_batch<64>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

extern "C" hipError_t msvk_launch_batch(int G, const BatArgs *a, int grid, int block, size_t lds, hipStream_t st) {
  if (grid < 1 || block < 64 || block > 1024 || block % 64 != 0 || lds > 160 * 1024) return hipErrorInvalidValue;
  switch (G) {
    case 8: hipLaunchKernelGGL(msv_batch<8>, dim3(grid), dim3(block), lds, st, *a); break;
    case 16: hipLaunchKernelGGL(msv_batch<16>, dim3(grid), dim3(block), lds, st, *a); break;
    case 32: hipLaunchKernelGGL(msv_batch<32>, dim3(grid), dim3(block), lds, st, *a); break;
    case 64: hipLaunchKernelGGL(msv_batch<64>, dim3(grid), dim3(block), lds, st, *a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Workgroups of msv_batch<G> resident per CU at this block size and LDS.
extern "C" hipError_t msvk_batch_occupancy(int G, int block, size_t lds, int *per_cu) {
  const void *f = G == 8 ? reinterpret_cast<const void *>(&msv_batch<8>)
                : G == 16 ? reinterpret_cast<const void *>(&msv_batch<16>)
                : G == 32 ? reinterpret_cast<const void *>(&msv_batch<32>)
                          : reinterpret_cast<const void *>(&msv_batch<64>);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, f, block, lds);
}

extern "C" hipError_t msvk_build_bat_table(const uint8_t *rbv, int M, const uint8_t *rowcode, int R, int G, int Q,
                                           int bias, uint32_t *out, hipStream_t st) {
  const int TW = G < 16 ? 16 : G;
  if ((G != 8 && G != 16 && G != 32) || Q < 1 || Q > (G == 16 ? 24 : 16) || R + 1 > 32)
    return hipErrorInvalidValue;
  const int64_t total = (int64_t) ((Q + 3) / 4) * 32 * TW * 4;
  int grid = (int) ((total + 255) / 256);
  if (grid > 65535) grid = 65535;
  hipLaunchKernelGGL(build_bat_table, dim3(grid), dim3(256), 0, st, rbv, M, rowcode, R, G, Q, TW, bias, out);
  return hipGetLastError();
}

// The grouped-scan kernels built: G = 8 up to 2 register blocks (M <= 128),
// 16 up to 6, 32 up to 4 and 64 up to 8.
#define MSVG_SCAN_GROUPS(X) \
  X(8, 1) X(8, 2) X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 5) X(16, 6) X(32, 1) X(32, 2) X(32, 3) X(32, 4) \
  X(64, 1) X(64, 2) X(64, 3) X(64, 4) X(64, 5) X(64, 6) X(64, 7) X(64, 8)

extern "C" int msvk_scan_group_built(int G, int QB) {
#define MSVG_SG_IS(g, qb) if (G == g && QB == qb) return 1;
  MSVG_SCAN_GROUPS(MSVG_SG_IS)
#undef MSVG_SG_IS
  return 0;
}

extern "C" hipError_t msvk_prepare_scan_group() {
  hipError_t e = hipSuccess;
#define MSVG_SG_ATTR(g, qb)                                                                                  \
  if (e == hipSuccess)                                                                                       \
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&msv_scan_group<g, qb>),                          \
                            hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  MSVG_SCAN_GROUPS(MSVG_SG_ATTR)
#undef MSVG_SG_ATTR
  return e;
}

extern "C" hipError_t msvk_launch_scan_group(int G, int QB, const ScanGroupArgs *a, int grid, int block, size_t lds,
                                             hipStream_t st) {
  if (grid < 1 || block < 64 || block > 1024 || block % 64 != 0 || lds > 160 * 1024 || a->np < 1 || grid % a->np != 0)
    return hipErrorInvalidValue;
#define MSVG_SG_LAUNCH(g, qb)                                                                                \
  if (G == g && QB == qb) {                                                                                  \
    hipLaunchKernelGGL((msv_scan_group<g, qb>), dim3(grid), dim3(block), lds, st, *a);                       \
    return hipGetLastError();                                                                                \
  }
  MSVG_SCAN_GROUPS(MSVG_SG_LAUNCH)
#undef MSVG_SG_LAUNCH
  return hipErrorInvalidValue;
}
