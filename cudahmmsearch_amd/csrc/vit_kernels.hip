// vit_kernels.hip -- gfx950 kernels for HMMER3's ViterbiFilter, the filter
// stage after MSV in p7_Pipeline (impl_sse/vitfilter.c:84-250), bit-exact.
//
// The reference runs Farrar's striped int16 recurrence over 8 words per SSE
// vector.  Here a target's model is striped over the int16 slots of G lanes
// (vit_band: G = 16 or 32 with several targets per wave, or 64) or of W
// waves (vit_kernel: W = 2, 4 or 8, models beyond 2048), two slots per VGPR:
// register j of lane c holds model positions k = j + s*Q + 1 for its slots
// s = 2c, 2c+1.  Every cell does the reference's operations in the
// reference's order with packed saturating int16 arithmetic (v_pk_add_i16
// clamp = _mm_adds_epi16, v_pk_max_i16 = _mm_max_epi16), so values do not
// depend on the striping.  Per row:
//   - the diagonal predecessors of register 0 come from the last register one
//     slot down: DPP wave_shr:1 (or row_shr:1) + v_alignbit, and across waves
//     through LDS;
//   - xE and Dmax (vitfilter.c:176, :198) are packed DPP max reductions, or
//     per-lane shares voted on against thresholds (vit_band);
//   - the special states run on the scalar unit;
//   - when the reference would run its lazy-F delete passes (Dmax + ddbound_w
//     > xB, :199) the kernel computes the same result, the closure
//     D(k) = max(D(k), D(k-1) + tDD(k-1)) along the model: a pass inside each
//     slot, then a max-plus scan across slots.  Saturating adds of costs
//     t <= 0 compose exactly (max(-32768, max(-32768, x+t1) + t2) =
//     max(-32768, x + t1 + t2)), so the scan carries unsaturated int32 sums.
// The score's float expression (:240-246) is evaluated on the host from xC.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include <utility>
#include "msv_device.h"

typedef short s16x2 __attribute__((ext_vector_type(2)));

namespace {

constexpr uint32_t NEG2 = 0x80008000u;   // (-32768, -32768): minus infinity in both halves

__device__ __forceinline__ s16x2 S(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ uint32_t U(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t adds(uint32_t a, uint32_t b) {
  return U(__builtin_elementwise_add_sat(S(a), S(b)));
}
__device__ __forceinline__ uint32_t maxs(uint32_t a, uint32_t b) { return U(__builtin_elementwise_max(S(a), S(b))); }
__device__ __forceinline__ int lo16(uint32_t v) { return (int) (int16_t) (v & 0xffffu); }
__device__ __forceinline__ int hi16(uint32_t v) { return (int) (int16_t) (v >> 16); }
__device__ __forceinline__ uint32_t pack16(int lo, int hi) {
  return (uint32_t) (uint16_t) lo | ((uint32_t) (uint16_t) hi << 16);
}
// x + T floored at -32768, for x >= -32768 and T <= 0 (a run of saturating adds).
__device__ __forceinline__ int fT(int x, int T) { return max(-32768, x + T); }

// Slot s <- slot s-1 for a register: the lane's low half takes the previous
// lane's high half, its high half its own low half; lane 0 takes the high
// half of edge (the last slot of the previous wave, or minus infinity).
__device__ __forceinline__ uint32_t shift_in(uint32_t last, uint32_t edge, int lane) {
  uint32_t prev = __builtin_amdgcn_mov_dpp(last, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
  prev = (lane == 0) ? edge : prev;
  return __builtin_amdgcn_alignbit(last, prev, 16);
}

// Packed int16 max over each row of 16 lanes, in every lane of the row: DPP
// rotations, so every lane has a source (max is idempotent, the wrap-around
// is harmless) and no fill value is needed.
__device__ __forceinline__ uint32_t row_max2(uint32_t v) {
  v = maxs(v, __builtin_amdgcn_mov_dpp(v, 0x121 /* row_ror:1 */, 0xf, 0xf, false));
  v = maxs(v, __builtin_amdgcn_mov_dpp(v, 0x122 /* row_ror:2 */, 0xf, 0xf, false));
  v = maxs(v, __builtin_amdgcn_mov_dpp(v, 0x124 /* row_ror:4 */, 0xf, 0xf, false));
  v = maxs(v, __builtin_amdgcn_mov_dpp(v, 0x128 /* row_ror:8 */, 0xf, 0xf, false));
  return v;
}

// Packed int16 max over the wave, returned in every lane (lane 63's value).
__device__ __forceinline__ uint32_t wave_max2(uint32_t v) {
  v = row_max2(v);
  v = maxs(v, __builtin_amdgcn_update_dpp(NEG2, v, 0x142 /* row_bcast:15 */, 0xa, 0xf, false));
  v = maxs(v, __builtin_amdgcn_update_dpp(NEG2, v, 0x143 /* row_bcast:31 */, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}


// The lane scan composes maps x -> max(c, x + T) (floor -32768), the source
// lane's map (from DPP control CTRL; rows outside ROW_MASK, and lanes whose
// source lies outside their row, read the identity (-32768, 0)) before this
// lane's: c <- max(c, c_src + T), T <- T_src + T.  The T terms are sums of
// tDD over slots, the same on every row, so they are scanned once per kernel
// (t_next) and each row's step only carries c (c_step).  No floor is needed
// there: c >= -32768 always, so max(c, max(-32768, c_src + T)) =
// max(c, c_src + T).  T is kept negated in 16 bits, saturated at 65535
// (negsat): a T <= -65535 acts as minus infinity either way, because
// c_src + T <= 32767 - 65535 = -32768 <= c.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int t_next(int T) {
  return T + __builtin_amdgcn_update_dpp(0, T, CTRL, ROW_MASK, 0xf, false);
}
__device__ __forceinline__ uint32_t negsat(int T) { return (uint32_t) min(-T, 65535); }
__device__ __forceinline__ uint32_t negsat2(int Ta, int Tb) { return negsat(Ta) | (negsat(Tb) << 16); }
// c <- max(c, c_src + T), T the HALF-th 16-bit field of tp, negated.
template <int CTRL, int ROW_MASK, int HALF>
__device__ __forceinline__ int c_step(int c, uint32_t tp) {
  const int co = __builtin_amdgcn_update_dpp(-32768, c, CTRL, ROW_MASK, 0xf, false);
  const int nt = HALF ? (int) (tp >> 16) : (int) (tp & 0xffffu);
  return max(c, co - nt);
}
// The same for the steps inside rows of 16 lanes (row_shr by 1, 2, 4, 8):
// a lane without a source reads 0, and its T term in tp is stored as 65535
// (minus infinity; in_row_t), so 0 - 65535 <= c leaves it unchanged.
template <int CTRL, int HALF>
__device__ __forceinline__ int c_step_row(int c, uint32_t tp) {
  const int co = __builtin_amdgcn_mov_dpp(c, CTRL, 0xf, 0xf, true);
  const int nt = HALF ? (int) (tp >> 16) : (int) (tp & 0xffffu);
  return max(c, co - nt);
}
// T of an in-row step with shift SH for this lane: minus infinity where the
// lane has no source in that step.
template <int SH>
__device__ __forceinline__ int in_row_t(int T, int lane) { return ((lane & 15) >= SH) ? T : -65535; }
// (max over both halves of xEv, max over both halves of Dmv) as one packed
// pair: two byte permutes and one v_pk_max_i16.
__device__ __forceinline__ uint32_t halves_max(uint32_t xEv, uint32_t Dmv) {
  const uint32_t lo = __builtin_amdgcn_perm(Dmv, xEv, 0x05040100u);   // (xEv.lo, Dmv.lo)
  const uint32_t hi = __builtin_amdgcn_perm(Dmv, xEv, 0x07060302u);   // (xEv.hi, Dmv.hi)
  return maxs(lo, hi);
}

}  // namespace

template <int Q, int W>
__global__ void __launch_bounds__(64 * W) vit_kernel(VitArgs a) {
  static_assert(W >= 2, "one wave per target runs vit_band");
  __shared__ uint32_t red[16], dlast[16], eM[16], eI[16], eD[16];
  __shared__ int scC[16], scT[16];
  constexpr int C = 64 * W;                            // lanes per target
  __shared__ uint4 tsc[C];                             // the lane scan's T terms per lane
  __shared__ uint32_t tse[C];                          // and its cross-wave terms
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int w = wave;
  const int c = w * 64 + lane;

  uint32_t tBM[Q], tMM[Q], tIM[Q], tDM[Q], tMD[Q], tMI[Q], tII[Q], tDD[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    tBM[j] = a.ttab[(0 * Q + j) * C + c];
    tMM[j] = a.ttab[(1 * Q + j) * C + c];
    tIM[j] = a.ttab[(2 * Q + j) * C + c];
    tDM[j] = a.ttab[(3 * Q + j) * C + c];
    tMD[j] = a.ttab[(4 * Q + j) * C + c];
    tMI[j] = a.ttab[(5 * Q + j) * C + c];
    tII[j] = a.ttab[(6 * Q + j) * C + c];
    tDD[j] = a.ttab[(7 * Q + j) * C + c];
  }
  int Tlo = 0, Thi = 0;                                // tDD summed over each slot
#pragma unroll
  for (int j = 0; j < Q; ++j) { Tlo += lo16(tDD[j]); Thi += hi16(tDD[j]); }
  // the lane scan's T terms, once: the lane's two slots, before each of the
  // six steps, after the last, and the previous lane's (exclusive) total
  const int Ts0 = Tlo + Thi;
  const int Ts1 = t_next<0x111, 0xf>(Ts0);
  const int Ts2 = t_next<0x112, 0xf>(Ts1);
  const int Ts3 = t_next<0x114, 0xf>(Ts2);
  const int Ts4 = t_next<0x118, 0xf>(Ts3);
  const int Ts5 = t_next<0x142, 0xa>(Ts4);
  const int Ts6 = t_next<0x143, 0xc>(Ts5);
  const int Tex = __builtin_amdgcn_update_dpp(0, Ts6, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  // kept in LDS, read on the rows that run the closure
  tsc[c] = make_uint4(negsat2(Tlo, Thi), negsat2(in_row_t<1>(Ts0, lane), in_row_t<2>(Ts1, lane)),
                      negsat2(in_row_t<4>(Ts2, lane), in_row_t<8>(Ts3, lane)), negsat2(Ts4, Ts5));
  tse[c] = negsat2(Ts6, Tex);
  __syncthreads();

  const uint32_t *et = a.etab;
  const int slot = (int) blockIdx.x;
  const int xN = a.base_w;                             // vitfilter.c:116; N loop cost is 0

  for (int it = a.uoff[slot], iend = a.uoff[slot + 1]; it < iend; ++it) {
    const int64_t s = a.items[it];
    const int L = a.db.len[s];
    const uint8_t *p = a.db.res + a.db.off[s];
    const int move = a.move[it];
    uint32_t Mr[Q], Ir[Q], Dr[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) Mr[j] = Ir[j] = Dr[j] = NEG2;        // :115
    int xB = (int16_t) (xN + move);                                  // :117
    int xJ = -32768, xC = -32768, status = 0;
    uint32_t edM = NEG2, edI = NEG2, edD = NEG2;
    // xB + tBM per register, the first term of every M cell: xB changes on
    // few rows (xJ is a running maximum), so it is recomputed only then.
    uint32_t bm[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xB, xB), tBM[j]);

    for (int i0 = 0; i0 < L && status == 0; i0 += 16) {
      const uint4 ch = *reinterpret_cast<const uint4 *>(p + i0);
      uint32_t w0 = ch.x, w1 = ch.y, w2 = ch.z, w3 = ch.w;   // 16 residues, consumed from the bottom
      const int nr = min(16, L - i0);
      for (int rr = 0; rr < nr; ++rr) {
        const int x = (int) (w0 & 0xffu);
        w0 = __builtin_amdgcn_alignbit(w1, w0, 8);
        w1 = __builtin_amdgcn_alignbit(w2, w1, 8);
        w2 = __builtin_amdgcn_alignbit(w3, w2, 8);
        w3 >>= 8;
        const uint32_t *ep = et + (size_t) x * Q * C + c;
        uint32_t e[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) e[j] = ep[j * C];
        uint32_t mpv = shift_in(Mr[Q - 1], edM, lane);               // :139-141
        uint32_t ipv = shift_in(Ir[Q - 1], edI, lane);
        uint32_t dpv = shift_in(Dr[Q - 1], edD, lane);
        uint32_t xEv = NEG2, Dmv = NEG2, dcv = NEG2;
        // Registers in descending order, each updated in place (no copies
        // across the row loop): register j reads register j-1's old M, I, D
        // (not yet updated) and its own old M, I; the D entering register
        // j+1 is this register's new M + tMD, written after register j+2
        // has read the old one.  The same operations per cell as :146-172.
#pragma unroll
        for (int j = Q - 1; j >= 0; --j) {
          const uint32_t pm = j ? Mr[j - 1] : mpv, pi = j ? Ir[j - 1] : ipv, pd = j ? Dr[j - 1] : dpv;
          uint32_t sv = bm[j];                                       // :146-150
          sv = maxs(sv, adds(pm, tMM[j]));
          sv = maxs(sv, adds(pi, tIM[j]));
          sv = maxs(sv, adds(pd, tDM[j]));
          sv = adds(sv, e[j]);
          xEv = maxs(xEv, sv);                                       // :151
          Ir[j] = maxs(adds(Mr[j], tMI[j]), adds(Ir[j], tII[j]));    // :171-172
          Mr[j] = sv;                                                // :161
          const uint32_t dn = adds(sv, tMD[j]);                      // :167
          Dmv = maxs(Dmv, dn);                                       // :168
          if (j < Q - 1) Dr[j + 1] = dn; else dcv = dn;              // :162
        }
        uint32_t r = wave_max2(halves_max(xEv, Dmv));
        uint32_t dedge = NEG2;
        {
          const uint32_t dl = __builtin_amdgcn_readlane(dcv, 63);
          if (lane == 0) { red[w] = r; dlast[w] = dl; }
          __syncthreads();
          r = red[0];
#pragma unroll
          for (int v = 1; v < W; ++v) r = maxs(r, red[v]);
          if (w > 0) dedge = dlast[w - 1];
        }
        const int xE = lo16(r), Dmax = hi16(r);                      // :176, :198
        if (xE >= 32767) { status = 16; break; }                     // :177
        xC = (int16_t) max(xC, xE + a.e_move);                       // :179 (C loop 0)
        xJ = (int16_t) max(xJ, xE + a.e_loop);                       // :180 (J loop 0)
        const int xBn = (int16_t) max(xJ + move, xN + move);         // :181
        if (xBn != xB) {
#pragma unroll
          for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xBn, xBn), tBM[j]);
        }
        xB = xBn;

        Dr[0] = shift_in(dcv, dedge, lane);                          // :230-231 / :203-204
        if (Dmax + a.ddbound_w > xB) {                               // :199
          const uint4 tt = tsc[c];
          // closure inside each slot
#pragma unroll
          for (int j = 1; j < Q; ++j) Dr[j] = maxs(Dr[j], adds(Dr[j - 1], tDD[j - 1]));
          const uint32_t outl = adds(Dr[Q - 1], tDD[Q - 1]);
          const int clo = lo16(outl);
          const uint32_t tsl = tt.x, ts01 = tt.y, ts23 = tt.z, ts45 = tt.w;
          // the lane's two slots as one map x -> max(c, x + T): low slot, then high
          int cs = max(hi16(outl), clo - (int) (tsl >> 16));        // clo through the high slot
          // inclusive scan over the lanes (lane l: lanes 0..l composed, in
          // order), in DPP: row_shr 1/2/4/8 inside each row of 16 lanes,
          // then row_bcast 15 and 31 across rows; lanes without a source
          // take the identity map (-32768, 0), which leaves them unchanged
          cs = c_step_row<0x111, 0>(cs, ts01);
          cs = c_step_row<0x112, 1>(cs, ts01);
          cs = c_step_row<0x114, 0>(cs, ts23);
          cs = c_step_row<0x118, 1>(cs, ts23);
          cs = c_step<0x142, 0xa, 0>(cs, ts45);
          cs = c_step<0x143, 0xc, 1>(cs, ts45);
          // exclusive: the previous lane's prefix (wave_shr:1; lane 0 the identity)
          const int ce = __builtin_amdgcn_update_dpp(-32768, cs, 0x138, 0xf, 0xf, false);
          int cin_lo = ce;
          {
            int carry = -32768;                                      // into this wave's first slot
            const uint32_t ts6e = tse[c];
            if (lane == 63) { scC[w] = cs; scT[w] = -(int) (ts6e & 0xffffu); }
            __syncthreads();
            for (int v = 0; v < w; ++v) carry = max(scC[v], fT(carry, scT[v]));
            cin_lo = max(ce, fT(carry, -(int) (ts6e >> 16)));
          }
          const int cin_hi = max(clo, cin_lo - (int) (tsl & 0xffffu));
          uint32_t y = pack16(cin_lo, cin_hi);
#pragma unroll
          for (int j = 0; j < Q; ++j) { Dr[j] = maxs(Dr[j], y); y = adds(y, tDD[j]); }
        }
        {
          if (lane == 63) { eM[w] = Mr[Q - 1]; eI[w] = Ir[Q - 1]; eD[w] = Dr[Q - 1]; }
          __syncthreads();
          if (w > 0) { edM = eM[w - 1]; edI = eI[w - 1]; edD = eD[w - 1]; }
        }
      }
    }
    // Several waves: every wave has read this item's last exchange (red[]
    // on a row that overflowed) before any wave's next item writes it.  The
    // item, its length and the overflow break are the same for all waves
    // of the workgroup, so all of them reach this barrier once per item.
    __syncthreads();
    if (threadIdx.x % C == 0) a.out[it] = (status << 16) | (int) (uint16_t) xC;
  }
}

// Workgroups of a kernel that fit one CU at once (at least 1).
template <class K>
static int resident_per_cu(K kernel, int block, size_t lds) {
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, lds) != hipSuccess || occ < 1) occ = 1;
  return occ;
}

// Grid of a multi-wave shape: one workgroup per target at a time, up to 8
// waves per CU.
template <int Q, int W>
static int grid_vit(const VitArgs *a, int ncu) {
  return std::max(1, std::min(a->n, ncu * std::max(1, 8 / W)));
}

template <int Q, int W>
static hipError_t launch_vit(const VitArgs *a, int ncu, hipStream_t st) {
  const int grid = grid_vit<Q, W>(a, ncu);
  hipLaunchKernelGGL((vit_kernel<Q, W>), dim3(grid), dim3(64 * W), 0, st, *a);
  return hipGetLastError();
}

template <int W, int... Qs>
static hipError_t dispatch_q(int Q, const VitArgs *a, int ncu, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t e = hipErrorInvalidValue;
  ((Q == Qs + 1 ? (e = launch_vit<Qs + 1, W>(a, ncu, st), 0) : 0), ...);
  return e;
}

// Q <= 16 registers per lane and W = 2, 4 or 8 waves per target (M up to
// 16384): the transitions of 16 registers still fit the 256 VGPRs a lane
// has when a workgroup holds at most two waves per SIMD.
extern "C" hipError_t msvk_launch_vit(int Q, int W, const VitArgs *a, int ncu, hipStream_t st) {
  if (Q < 1 || Q > 16 || a->n < 1) return hipErrorInvalidValue;
  auto qs = std::make_integer_sequence<int, 16>{};
  switch (W) {
    case 2: return dispatch_q<2>(Q, a, ncu, st, qs);
    case 4: return dispatch_q<4>(Q, a, ncu, st, qs);
    case 8: return dispatch_q<8>(Q, a, ncu, st, qs);
    default: return hipErrorInvalidValue;
  }
}

// The slots of a multi-wave launch: its workgroups, each running the targets
// the host lists for it in a.uoff.
template <int... Qs>
static int slots_vit(int Q, int W, const VitArgs *a, int ncu, std::integer_sequence<int, Qs...>) {
  int r = 0;
  ((Q == Qs + 1 ? (r = (W == 2) ? grid_vit<Qs + 1, 2>(a, ncu)
                    : (W == 4) ? grid_vit<Qs + 1, 4>(a, ncu) : grid_vit<Qs + 1, 8>(a, ncu), 0) : 0), ...);
  return r;
}

extern "C" int msvk_vit_slots(int Q, int W, const VitArgs *a, int ncu) {
  if (a->n < 1 || Q < 1 || Q > 16 || (W != 2 && W != 4 && W != 8)) return 0;
  return slots_vit(Q, W, a, ncu, std::make_integer_sequence<int, 16>{});
}

namespace {

// Packed int16 max over each G-lane group; the group's last lane holds it.
template <int G>
__device__ __forceinline__ uint32_t group_max2(uint32_t v) {
  v = row_max2(v);
  if constexpr (G >= 32) v = maxs(v, __builtin_amdgcn_update_dpp(NEG2, v, 0x142 /* row_bcast:15 */, 0xa, 0xf, false));
  if constexpr (G == 64) v = maxs(v, __builtin_amdgcn_update_dpp(NEG2, v, 0x143 /* row_bcast:31 */, 0xc, 0xf, false));
  return v;
}

// Per-lane value of a per-group scalar: the lanes of group g take v[g].
// ge[gg] is all ones in the lanes of groups >= gg and zero elsewhere, made
// opaque to the compiler once per kernel, so each step is one v_bfi_b32 (a
// select chain on the lane's group index would be turned into an indexed
// array in scratch memory).
template <int NG>
__device__ __forceinline__ int of_group(const int (&v)[NG], const int (&ge)[NG]) {
  int r = v[0];
#pragma unroll
  for (int gg = 1; gg < NG; ++gg) r = (ge[gg] & v[gg]) | (~ge[gg] & r);
  return r;
}

// Ballot bits of one G-lane group.
template <int G>
constexpr uint64_t GMASK = (G == 64) ? ~0ull : ((1ull << G) - 1);

}  // namespace

// ---------------------------------------------------------------------------
// vit_band<G, Q, VOTE>: models up to 2560 positions, Q <= 16 registers per
// lane (20 for G = 64), NG = 64/G targets per wave with G = 16, 32 or 64 lanes each (the
// wave's targets share the per-row fixed work: the shifts, the votes, the
// special states, the delete-closure scan).  Per row:
//   - rows run four at a time, so a residue is one byte extract;
//   - the next row's emissions are read from LDS while this row's special
//     states and delete closure run;
//   - the slot shifts keep their fill value (minus infinity at a group's
//     first lane) in a register the shift never writes there, so a shift is
//     one DPP move and one v_alignbit;
//   - xE and Dmax are not reduced on every row: each lane keeps its share
//     (the max over its registers) and the wave votes (ballot) against two
//     thresholds per group, tE = min(max(xJ, xN) - e_loop, 32766) and
//     tD = xB - ddbound_w.  Only a row where some share exceeds tE can
//     overflow or raise xJ (and with it xB), so only then is xE reduced and
//     the special states updated (vitfilter.c:176-181); xC, a running
//     maximum nothing else reads, is each lane's largest share, reduced when
//     the targets end;
//   - when some group's lazy-F test holds (vitfilter.c:199) -- or on every
//     row, without the Dmax shares and the vote, for profiles whose test
//     holds on most rows (!VOTE) -- the whole wave runs the delete closure:
//     the pass inside each slot, the max-plus scan over the slots with its T
//     terms in registers, the carries applied.  The closure is applied to
//     every group: for a group whose test fails it changes no M value,
//     because every D->...->D->M path it adds is bounded by
//     Dmax + ddbound_w + tBM <= xB + tBM, the B->M term of the same cell
//     (the bound the reference's test rests on, vitfilter.c:183-197), and D
//     itself enters later rows only through M;
//   - a group whose target has ended (or overflowed) keeps computing padding
//     rows but no longer votes or updates its states.
// The emission table (G lanes wide) is staged in LDS; each wave runs the
// units (NG targets) the host lists for it.
// ---------------------------------------------------------------------------
namespace {

// Slot s <- slot s-1 inside each G-lane group.  keep holds minus infinity in
// every group's first lane: G = 16 lanes are DPP rows, whose first lane the
// row shift never writes, and lane 0 has no source in the wave shift; G = 32
// selects it back for the second group.
template <int G>
__device__ __forceinline__ uint32_t shift_keep(uint32_t last, uint32_t &keep, int lane) {
  if constexpr (G == 16) {
    keep = __builtin_amdgcn_update_dpp(keep, last, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
  } else if constexpr (G == 64) {
    keep = __builtin_amdgcn_update_dpp(keep, last, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
  } else {
    const uint32_t prev = __builtin_amdgcn_update_dpp(keep, last, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    keep = (lane == 32) ? NEG2 : prev;
  }
  return __builtin_amdgcn_alignbit(last, keep, 16);
}

// One in-row scan step with its negated T term nt in a register: c <- max(c,
// c_src - nt); a lane without a source reads 0 and holds nt = 65535.
template <int CTRL>
__device__ __forceinline__ int c_row(int c, int nt) {
  return max(c, __builtin_amdgcn_mov_dpp(c, CTRL, 0xf, 0xf, true) - nt);
}

// The scan's negated T terms of one lane (see t_next, negsat, in_row_t).
struct ScanT {
  int lo, hi;            // the lane's low and high slot
  int s1, s2, s4, s8;    // in-row steps
  int s16;               // row_bcast:15 step (G >= 32)
  int s32;               // row_bcast:31 step (G = 64)
};

// The delete closure over all lanes (vitfilter.c:203-227; see close_d).
template <int G, int Q>
__device__ __forceinline__ void close_all(uint32_t (&Dr)[Q], const uint32_t (&tDD)[Q], const ScanT &t, int &kC,
                                          int lane) {
#pragma unroll
  for (int j = 1; j < Q; ++j) Dr[j] = maxs(Dr[j], adds(Dr[j - 1], tDD[j - 1]));
  const uint32_t outl = adds(Dr[Q - 1], tDD[Q - 1]);
  const int clo = lo16(outl);
  int cs = max(hi16(outl), clo - t.hi);              // the lane's two slots as one map
  cs = c_row<0x111>(cs, t.s1);
  cs = c_row<0x112>(cs, t.s2);
  cs = c_row<0x114>(cs, t.s4);
  cs = c_row<0x118>(cs, t.s8);
  int ce;
  if constexpr (G >= 32)
    cs = max(cs, __builtin_amdgcn_update_dpp(-32768, cs, 0x142 /* row_bcast:15 */, 0xa, 0xf, false) - t.s16);
  if constexpr (G == 64) {
    cs = max(cs, __builtin_amdgcn_update_dpp(-32768, cs, 0x143 /* row_bcast:31 */, 0xc, 0xf, false) - t.s32);
    // exclusive: the previous lane's prefix; lane 0 keeps -32768
    kC = __builtin_amdgcn_update_dpp(kC, cs, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    ce = kC;
  } else if constexpr (G == 32) {
    // exclusive: the previous lane's prefix; a group's first lane -32768
    const int prev = __builtin_amdgcn_update_dpp(kC, cs, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
    kC = (lane == 32) ? -32768 : prev;
    ce = kC;
  } else {
    kC = __builtin_amdgcn_update_dpp(kC, cs, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
    ce = kC;
  }
  const int cin_hi = max(clo, ce - t.lo);
  uint32_t y = __builtin_amdgcn_perm((uint32_t) cin_hi, (uint32_t) ce, 0x05040100u);   // (ce, cin_hi)
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    Dr[j] = maxs(Dr[j], y);
    if (j < Q - 1) y = adds(y, tDD[j]);
  }
}

}  // namespace

// Workgroup size: 64-lane targets, whose tables are the largest, share one
// LDS copy among eight waves.
template <int G> constexpr int band_block() { return G == 64 ? 512 : 256; }
// Models beyond 2048 positions (G = 64, Q = 17 to 20) read their MM, IM and
// DM transitions from LDS each row, which keeps the lane within 256 VGPRs
// with two waves per SIMD.
template <int Q> constexpr bool band_lt() { return Q > 16; }

template <int G, int Q, bool VOTE>
__global__ void __launch_bounds__(band_block<G>()) __attribute__((amdgpu_waves_per_eu(Q <= 5 ? 4 : (Q <= 8 ? 3 : 2), 8)))
vit_band(VitArgs a) {
  extern __shared__ uint32_t etab_lds[];
  constexpr int NG = 64 / G;
  const VitProfDev pd = a.profs[a.wg_prof[blockIdx.x]];   // this workgroup's profile
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int c = lane & (G - 1);                        // lane within its group
  const int g = lane / G;                              // the wave's target this lane works on
  int ge[NG];
#pragma unroll
  for (int gg = 0; gg < NG; ++gg) {
    ge[gg] = (g >= gg) ? -1 : 0;
    asm volatile("" : "+v"(ge[gg]));
  }

  constexpr bool LT = band_lt<Q>();
  constexpr int QR = LT ? 1 : Q;
  uint32_t tMM[QR], tIM[QR], tDM[QR], tMD[Q], tMI[Q], tII[Q], tDD[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    if constexpr (!LT) {
      tMM[j] = pd.ttab[(1 * Q + j) * G + c];
      tIM[j] = pd.ttab[(2 * Q + j) * G + c];
      tDM[j] = pd.ttab[(3 * Q + j) * G + c];
    }
    tMD[j] = pd.ttab[(4 * Q + j) * G + c];
    tMI[j] = pd.ttab[(5 * Q + j) * G + c];
    tII[j] = pd.ttab[(6 * Q + j) * G + c];
    tDD[j] = pd.ttab[(7 * Q + j) * G + c];
  }
  int Tlo = 0, Thi = 0;                                // tDD summed over each slot
#pragma unroll
  for (int j = 0; j < Q; ++j) { Tlo += lo16(tDD[j]); Thi += hi16(tDD[j]); }
  ScanT st_;
  {
    const int Ts0 = Tlo + Thi;
    const int Ts1 = t_next<0x111, 0xf>(Ts0);
    const int Ts2 = t_next<0x112, 0xf>(Ts1);
    const int Ts3 = t_next<0x114, 0xf>(Ts2);
    const int Ts4 = (G >= 32) ? t_next<0x118, 0xf>(Ts3) : 0;
    const int Ts5 = (G == 64) ? t_next<0x142, 0xa>(Ts4) : 0;
    st_.lo = (int) negsat(Tlo);
    st_.hi = (int) negsat(Thi);
    st_.s1 = (int) negsat(in_row_t<1>(Ts0, lane));
    st_.s2 = (int) negsat(in_row_t<2>(Ts1, lane));
    st_.s4 = (int) negsat(in_row_t<4>(Ts2, lane));
    st_.s8 = (int) negsat(in_row_t<8>(Ts3, lane));
    st_.s16 = (int) negsat(Ts4);
    st_.s32 = (int) negsat(Ts5);
  }

  const int nemit = (a.db.R + 1) * Q * G;
  for (int i = threadIdx.x; i < nemit; i += blockDim.x) etab_lds[i] = pd.etab[i];
  if constexpr (LT)
    for (int i = threadIdx.x; i < 3 * Q * G; i += blockDim.x) etab_lds[nemit + i] = pd.ttab[Q * G + i];
  __syncthreads();
  const uint32_t *tl = etab_lds + nemit + c;           // LT: this lane's MM, IM, DM

  const uint32_t nb = (uint32_t) a.db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);
  const int xN = pd.base_w;                             // vitfilter.c:116; N loop cost is 0
  const int slot = blockIdx.x * (blockDim.x >> 6) + wave;

  for (int k = a.uoff[slot], kend = a.uoff[slot + 1]; k < kend; ++k) {
    const int it = a.ustart[k] + g;
    const bool valid = it < pd.end;
    const int64_t s = valid ? a.items[it] : 0;
    const int L = valid ? a.db.len[s] : 0;
    const int Lpad = (L + 15) & ~15;
    const uint8_t *p = a.db.res + (valid ? a.db.off[s] : 0);
    const int mvl = valid ? (int) a.move[it] : 0;

    int mv[NG], xB[NG], xJ[NG], st[NG];
    int Lmax = 0, Lmin = 0x7fffffff;
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      const int Lg = __builtin_amdgcn_readlane(L, gg * G);
      Lmax = max(Lmax, Lg);
      Lmin = min(Lmin, Lg);
      mv[gg] = __builtin_amdgcn_readlane(mvl, gg * G);
      xB[gg] = (int16_t) (xN + mv[gg]);                              // :117
      xJ[gg] = -32768; st[gg] = 0;
    }
    uint32_t Mr[Q], Ir[Q], Dr[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) Mr[j] = Ir[j] = Dr[j] = NEG2;        // :115
    uint32_t bm[Q];
    int tE, tD;
    auto refresh = [&]() {
      int te[NG], td[NG];
#pragma unroll
      for (int gg = 0; gg < NG; ++gg) {
        te[gg] = st[gg] ? 0x7fffffff : min(max(xJ[gg], xN) - pd.e_loop, 32766);
        td[gg] = xB[gg] - pd.ddbound_w;
      }
      tE = of_group<NG>(te, ge);
      tD = of_group<NG>(td, ge);
      const int xb = of_group<NG>(xB, ge);
#pragma unroll
      for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xb, xb), pd.ttab[j * G + c]);
    };
    refresh();
    uint32_t xacc = NEG2;               // low half: this lane's largest row maximum (xC at the end)
    uint32_t kM = NEG2, kI = NEG2, kD = NEG2, kD0 = NEG2;   // shift fills (see shift_keep)
    int kC = -32768;                    // the closure scan's exclusive fill
    // The emissions of the next row are read from LDS while the current
    // row's special states and delete closure run (held behind compiler
    // fences so the reads stay there).
    uint32_t e[Q];
    auto load_e = [&](int x) {
      const uint32_t *ep = etab_lds + x * (Q * G) + c;
#pragma unroll
      for (int j = 0; j < Q; ++j) e[j] = ep[j * G];
    };

    // One row; all_alive: every group's target has this row (no masks).
    auto row = [&](int xnext, bool all_alive, bool alive) {
      const uint32_t mpv = shift_keep<G>(Mr[Q - 1], kM, lane);      // :139-141
      const uint32_t ipv = shift_keep<G>(Ir[Q - 1], kI, lane);
      const uint32_t dpv = shift_keep<G>(Dr[Q - 1], kD, lane);
      uint32_t xEv = NEG2, Dmv = NEG2, dcv = NEG2;
      // Registers in descending order, each updated in place (see vit_group).
#pragma unroll
      for (int j = Q - 1; j >= 0; --j) {
        const uint32_t pm = j ? Mr[j - 1] : mpv, pi = j ? Ir[j - 1] : ipv, pd = j ? Dr[j - 1] : dpv;
        const uint32_t mm = LT ? tl[(0 * Q + j) * G] : tMM[LT ? 0 : j];
        const uint32_t im = LT ? tl[(1 * Q + j) * G] : tIM[LT ? 0 : j];
        const uint32_t dm = LT ? tl[(2 * Q + j) * G] : tDM[LT ? 0 : j];
        // :146-150 as a tree (max is exact in any order): two independent
        // maxima, so dependent packed ops need not follow each other
        uint32_t sv = maxs(maxs(bm[j], adds(pm, mm)), maxs(adds(pi, im), adds(pd, dm)));
        sv = adds(sv, e[j]);
        xEv = maxs(xEv, sv);                                         // :151
        Ir[j] = maxs(adds(Mr[j], tMI[j]), adds(Ir[j], tII[j]));      // :171-172
        Mr[j] = sv;                                                  // :161
        const uint32_t dn = adds(sv, tMD[j]);                        // :167
        if constexpr (VOTE) Dmv = maxs(Dmv, dn);                     // :168
        if (j < Q - 1) Dr[j + 1] = dn; else dcv = dn;                // :162
      }
      asm volatile("" ::: "memory");
      load_e(xnext);
      asm volatile("" ::: "memory");
      const uint32_t hv = halves_max(xEv, Dmv);                      // (xE share, Dmax share)
      // Votes against per-lane thresholds; a lane past its target's end
      // takes thresholds no value reaches.
      const int te = (all_alive || alive) ? tE : 0x7fffffff;
      const int td = (all_alive || alive) ? tD : 0x7fffffff;
      xacc = maxs(xacc, (all_alive || alive) ? hv : NEG2);
      const uint64_t mE = __builtin_amdgcn_ballot_w64(lo16(hv) > te);
      if (mE) {
        const uint32_t r = group_max2<G>(hv);
#pragma unroll
        for (int gg = 0; gg < NG; ++gg) {
          // every group's states, unchanged where its lanes did not vote
          const bool hit = ((mE >> (gg * G)) & GMASK<G>) != 0;
          const int xE = lo16(__builtin_amdgcn_readlane(r, gg * G + G - 1));
          const bool over = hit && xE >= 32767;                             // :177
          const bool upd = hit && !over;
          st[gg] = over ? 16 : st[gg];
          xJ[gg] = upd ? (int) (int16_t) max(xJ[gg], xE + pd.e_loop) : xJ[gg];   // :180 (J loop 0)
          xB[gg] = upd ? (int) (int16_t) max(xJ[gg] + mv[gg], xN + mv[gg]) : xB[gg];   // :181
        }
        refresh();
      }
      Dr[0] = shift_keep<G>(dcv, kD0, lane);                         // :230-231 / :203-204
      // :199 for some group (VOTE), or on every row: profiles whose test
      // holds on most rows skip the Dmax share and the vote
      if (!VOTE || __builtin_amdgcn_ballot_w64(hi16(hv) > td))
        close_all<G, Q>(Dr, tDD, st_, kC, lane);                     // :203-227
    };

    // Chunks of 16 residues, read two chunks ahead; the loaded value is
    // replaced by the null row only where it is used.
    const uint4 c0 = *reinterpret_cast<const uint4 *>(p);
    const uint4 c1 = *reinterpret_cast<const uint4 *>(p + (16 < Lpad ? 16 : 0));
    uint4 cur = (0 < Lpad) ? c0 : null4;
    uint4 nxt = (16 < Lpad) ? c1 : null4;
    load_e((int) (cur.x & 0xffu));
    for (int i0 = 0; i0 < Lmax; i0 += 16) {
      const int i2 = i0 + 32;
      const uint4 raw = *reinterpret_cast<const uint4 *>(p + (i2 < Lpad ? i2 : 0));
      uint32_t w0 = cur.x, w1 = cur.y, w2 = cur.z, w3 = cur.w;   // 16 residues, four per word
      // Rows past Lmax up to the next multiple of four are computed on the
      // residues the padded chunk holds; no group is alive on them.
      const int nq = min(4, (Lmax - i0 + 3) >> 2);
      for (int qd = 0; qd < nq; ++qd) {
        const int ib = i0 + 4 * qd;
#pragma unroll
        for (int r = 0; r < 4; ++r)
          row((int) (((r < 3 ? w0 >> (8 * r + 8) : w1)) & 0xffu), false, ib + r < L);
        w0 = w1; w1 = w2; w2 = w3; w3 = nxt.x;
      }
      cur = nxt;
      nxt = (i2 < Lpad) ? raw : null4;
    }
    // xC = max over the target's rows of xE + e_move (:179, C loop 0)
    const uint32_t xm = group_max2<G>(xacc);
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      const int xC = max(-32768, lo16(__builtin_amdgcn_readlane(xm, gg * G + G - 1)) + pd.e_move);
      if (lane == gg * G && valid) a.out[it] = (st[gg] << 16) | (int) (uint16_t) xC;
    }
  }
}

template <int G, int Q>
static size_t band_lds(int R) {
  return ((size_t) (R + 1) * Q * G + (band_lt<Q>() ? 3 * Q * G : 0)) * 4;
}

template <int G, int Q, bool VOTE>
static int cap_band(int R, int ncu) {
  return ncu * resident_per_cu(vit_band<G, Q, VOTE>, band_block<G>(), band_lds<G, Q>(R));
}

template <int G, int Q, bool VOTE>
static hipError_t launch_band(const VitArgs *a, int grid, hipStream_t st) {
  const size_t lds = band_lds<G, Q>(a->db.R);
  hipLaunchKernelGGL((vit_band<G, Q, VOTE>), dim3(grid), dim3(band_block<G>()), lds, st, *a);
  return hipGetLastError();
}

template <int G, bool VOTE, int... Qs>
static hipError_t dispatch_band(int Q, const VitArgs *a, int grid, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t e = hipErrorInvalidValue;
  ((Q == Qs + 1 ? (e = launch_band<G, Qs + 1, VOTE>(a, grid, st), 0) : 0), ...);
  return e;
}

template <int G, bool VOTE, int... Qs>
static int caps_band(int Q, int R, int ncu, std::integer_sequence<int, Qs...>) {
  int r = 0;
  ((Q == Qs + 1 ? (r = cap_band<G, Qs + 1, VOTE>(R, ncu), 0) : 0), ...);
  return r;
}

// G = 16, 32 or 64 lanes per target, Q <= 16 registers per lane (20 for
// G = 64); vote: the rows vote on the lazy-F test (else every row closes the
// delete paths).  grid workgroups of msvk_vit_band_block(G) threads.
extern "C" hipError_t msvk_launch_vit_band(int G, int Q, int vote, const VitArgs *a, int grid, hipStream_t st) {
  if (Q < 1 || Q > (G == 64 ? 20 : 16) || grid < 1) return hipErrorInvalidValue;
  auto qs = std::make_integer_sequence<int, 16>{};
  auto qw = std::make_integer_sequence<int, 20>{};
  if (G == 16) return vote ? dispatch_band<16, true>(Q, a, grid, st, qs) : dispatch_band<16, false>(Q, a, grid, st, qs);
  if (G == 32) return vote ? dispatch_band<32, true>(Q, a, grid, st, qs) : dispatch_band<32, false>(Q, a, grid, st, qs);
  if (G == 64) return vote ? dispatch_band<64, true>(Q, a, grid, st, qw) : dispatch_band<64, false>(Q, a, grid, st, qw);
  return hipErrorInvalidValue;
}

// Threads per workgroup of a vit_band shape.
extern "C" int msvk_vit_band_block(int G) { return G == 64 ? band_block<64>() : (G == 32 ? band_block<32>() : band_block<16>()); }

// Workgroups of a vit_band shape that the device holds at once for a
// database of R + 1 rows (0 if the shape does not exist).
extern "C" int msvk_vit_band_cap(int G, int Q, int vote, int R, int ncu) {
  if (Q < 1 || Q > (G == 64 ? 20 : 16) || (G != 16 && G != 32 && G != 64)) return 0;
  auto qs = std::make_integer_sequence<int, 16>{};
  auto qw = std::make_integer_sequence<int, 20>{};
  if (G == 16) return vote ? caps_band<16, true>(Q, R, ncu, qs) : caps_band<16, false>(Q, R, ncu, qs);
  if (G == 32) return vote ? caps_band<32, true>(Q, R, ncu, qs) : caps_band<32, false>(Q, R, ncu, qs);
  return vote ? caps_band<64, true>(Q, R, ncu, qw) : caps_band<64, false>(Q, R, ncu, qw);
}

// ---------------------------------------------------------------------------
// vit_seg<Q>: models beyond one wave's tables, cut into column segments of
// 128 Q positions; one launch per segment over a profile's items, one
// target per wave (vit_band<64, Q>'s row, eight waves sharing the segment's
// tables in LDS, the delete paths closed on every row).  Row i of a segment
// takes, where the segment's first position needs them,
//   - the M, I and D values of row i-1 at the previous segment's last
//     position (the diagonal into position k0 + 1), and
//   - the D value entering position k0 + 1 in row i (max of M + tMD and
//     D + tDD at the previous segment's last position, D after its closure),
// and leaves its own for the next segment.  Each of the four moves through
// one register holding two queues (the msv_seg scheme): lanes 0..15 the
// incoming values of a block of 16 rows (lane 0 the current row's), lanes
// 48..63 the outgoing ones; one row_shl:1 per row advances both, the row's
// shift takes lane 0 where it leaves a hole, lane 63 takes the new value.
// B stays xN + move on every row (the J-free case): a row where some lane's
// xE share exceeds min(xN - e_loop, 32766) could raise J above xN (and with
// it B) or overflow, so such an item is marked and, after the last segment,
// scored again by the exact kernel (VIT_DEFER).  Otherwise J never exceeds
// xN, B never changes, and xC = max over rows of xE + e_move (C loop 0), the
// row maxima being maxima over the segments' shares: the result is exact.
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(Q <= 5 ? 4 : (Q <= 8 ? 3 : 2), 8)))
vit_seg(VitSegArgs a) {
  extern __shared__ uint32_t etab_lds[];
  constexpr int G = 64;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int c = lane;

  constexpr bool LT = band_lt<Q>();
  constexpr int QR = LT ? 1 : Q;
  uint32_t tMM[QR], tIM[QR], tDM[QR], tMD[Q], tMI[Q], tII[Q], tDD[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    if constexpr (!LT) {
      tMM[j] = a.ttab[(1 * Q + j) * G + c];
      tIM[j] = a.ttab[(2 * Q + j) * G + c];
      tDM[j] = a.ttab[(3 * Q + j) * G + c];
    }
    tMD[j] = a.ttab[(4 * Q + j) * G + c];
    tMI[j] = a.ttab[(5 * Q + j) * G + c];
    tII[j] = a.ttab[(6 * Q + j) * G + c];
    tDD[j] = a.ttab[(7 * Q + j) * G + c];
  }
  int Tlo = 0, Thi = 0;
#pragma unroll
  for (int j = 0; j < Q; ++j) { Tlo += lo16(tDD[j]); Thi += hi16(tDD[j]); }
  ScanT st_;
  {
    const int Ts0 = Tlo + Thi;
    const int Ts1 = t_next<0x111, 0xf>(Ts0);
    const int Ts2 = t_next<0x112, 0xf>(Ts1);
    const int Ts3 = t_next<0x114, 0xf>(Ts2);
    const int Ts4 = t_next<0x118, 0xf>(Ts3);
    const int Ts5 = t_next<0x142, 0xa>(Ts4);
    st_.lo = (int) negsat(Tlo);
    st_.hi = (int) negsat(Thi);
    st_.s1 = (int) negsat(in_row_t<1>(Ts0, lane));
    st_.s2 = (int) negsat(in_row_t<2>(Ts1, lane));
    st_.s4 = (int) negsat(in_row_t<4>(Ts2, lane));
    st_.s8 = (int) negsat(in_row_t<8>(Ts3, lane));
    st_.s16 = (int) negsat(Ts4);
    st_.s32 = (int) negsat(Ts5);
  }

  const int nemit = (a.db.R + 1) * Q * G;
  for (int i = threadIdx.x; i < nemit; i += blockDim.x) etab_lds[i] = a.etab[i];
  if constexpr (LT)
    for (int i = threadIdx.x; i < 3 * Q * G; i += blockDim.x) etab_lds[nemit + i] = a.ttab[Q * G + i];
  __syncthreads();
  const uint32_t *tl = etab_lds + nemit + c;           // LT: this lane's MM, IM, DM

  const uint32_t nb = (uint32_t) a.db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);
  const int xN = a.base_w;                              // vitfilter.c:116; N loop cost is 0
  const int tE = min(xN - a.e_loop, 32766);             // xJ = -32768 < xN: max(xJ, xN) = xN
  constexpr uint32_t NEGH = 0x80000000u;                // -32768 in the high half
  const int slot = blockIdx.x * (blockDim.x >> 6) + wave;

  for (int it = a.uoff[slot], iend = a.uoff[slot + 1]; it < iend; ++it) {
    const int64_t s = a.items[it];
    const int L = a.db.len[s];
    const int Lpad = (L + 15) & ~15;
    const uint8_t *p = a.db.res + a.db.off[s];
    const int64_t co = a.coff[it];
    int16_t *cM = a.cM + co, *cI = a.cI + co, *cD = a.cD + co, *cE = a.cE + co;
    const int xB = (int16_t) (xN + (int) a.move[it]);                 // :117, constant here
    uint32_t bm[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xB, xB), a.ttab[j * G + c]);
    uint32_t Mr[Q], Ir[Q], Dr[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) Mr[j] = Ir[j] = Dr[j] = NEG2;        // :115
    uint32_t xacc = NEG2;                // low half: this lane's largest xE share
    int kC = -32768;                     // the closure scan's exclusive fill
    int fired = 0;
    uint32_t qM = NEGH, qI = NEGH, qD = NEGH, qE = NEGH;     // carry queues
    uint32_t clM = NEGH, clI = NEGH, clD = NEGH;             // carries of the row before the block
    uint32_t dent = NEG2;                // the D value leaving the segment in the row before
    uint32_t e[Q];
    auto load_e = [&](int x) {
      const uint32_t *ep = etab_lds + x * (Q * G) + c;
#pragma unroll
      for (int j = 0; j < Q; ++j) e[j] = ep[j * G];
    };
    // a queue step: the row's shifted register, lane 0 from the queue
    auto qshift = [](uint32_t last, uint32_t &q, uint32_t out_prev) -> uint32_t {
      const uint32_t xs = __builtin_amdgcn_update_dpp(out_prev, q, 0x101 /* row_shl:1 */, 0xf, 0xf, false);
      const uint32_t prev = __builtin_amdgcn_update_dpp(q, last, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
      q = xs;
      return __builtin_amdgcn_alignbit(last, prev, 16);
    };

    auto row = [&](int xnext, bool alive) {
      const uint32_t lM = Mr[Q - 1], lI = Ir[Q - 1], lD = Dr[Q - 1];
      const uint32_t mpv = qshift(lM, qM, lM);                      // :139-141
      const uint32_t ipv = qshift(lI, qI, lI);
      const uint32_t dpv = qshift(lD, qD, lD);
      uint32_t xEv = NEG2, dcv = NEG2;
#pragma unroll
      for (int j = Q - 1; j >= 0; --j) {
        const uint32_t pm = j ? Mr[j - 1] : mpv, pi = j ? Ir[j - 1] : ipv, pd = j ? Dr[j - 1] : dpv;
        const uint32_t mm = LT ? tl[(0 * Q + j) * G] : tMM[LT ? 0 : j];
        const uint32_t im = LT ? tl[(1 * Q + j) * G] : tIM[LT ? 0 : j];
        const uint32_t dm = LT ? tl[(2 * Q + j) * G] : tDM[LT ? 0 : j];
        // :146-150 as a tree (max is exact in any order): two independent
        // maxima, so dependent packed ops need not follow each other
        uint32_t sv = maxs(maxs(bm[j], adds(pm, mm)), maxs(adds(pi, im), adds(pd, dm)));
        sv = adds(sv, e[j]);
        xEv = maxs(xEv, sv);                                         // :151
        Ir[j] = maxs(adds(Mr[j], tMI[j]), adds(Ir[j], tII[j]));      // :171-172
        Mr[j] = sv;                                                  // :161
        const uint32_t dn = adds(sv, tMD[j]);                        // :167
        if (j < Q - 1) Dr[j + 1] = dn; else dcv = dn;                // :162
      }
      asm volatile("" ::: "memory");
      load_e(xnext);
      asm volatile("" ::: "memory");
      const uint32_t hv = halves_max(xEv, NEG2);                     // low half: xE share
      xacc = maxs(xacc, alive ? hv : NEG2);
      if (__builtin_amdgcn_ballot_w64(lo16(hv) > (alive ? tE : 0x7fffffff))) fired = 1;
      Dr[0] = qshift(dcv, qE, dent);                                 // :230-231 / :203-204
      close_all<G, Q>(Dr, tDD, st_, kC, lane);                       // :203-227
      dent = maxs(dcv, adds(Dr[Q - 1], tDD[Q - 1]));
    };

    // Carry loads are issued a block ahead (rows i .. i+15 in lanes 0..15,
    // high halves), and a block's outgoing carries are stored when the next
    // block starts, so no block waits for a load or store it has just
    // issued.  A block writes only its own rows, which the loads of the
    // next block do not touch.
    uint32_t pM = NEGH, pI = NEGH, pD = NEGH, pE = NEGH;
    auto load_carries = [&](int i) {
      pM = pI = pD = pE = NEGH;
      if (!a.first && lane < 16 && i < Lpad) {
        pM = (uint32_t) (uint16_t) cM[i + lane] << 16;
        pI = (uint32_t) (uint16_t) cI[i + lane] << 16;
        pD = (uint32_t) (uint16_t) cD[i + lane] << 16;
        pE = (uint32_t) (uint16_t) cE[i + lane] << 16;
      }
    };
    uint32_t sM = 0, sI = 0, sD = 0, sE = 0;             // outgoing carries not yet stored
    int snr = 0, si = 0;                                  // their rows, from row si
    auto store_carries = [&]() {
      if (!a.last && snr > 0 && lane >= 64 - snr) {
        const int pos = si + lane - (64 - snr);            // lane 64 - snr + k holds row si + k
        cM[pos] = (int16_t) (sM >> 16);
        cI[pos] = (int16_t) (sI >> 16);
        cD[pos] = (int16_t) (sD >> 16);
        cE[pos] = (int16_t) (sE >> 16);
      }
    };

    const uint4 c0 = *reinterpret_cast<const uint4 *>(p);
    const uint4 c1 = *reinterpret_cast<const uint4 *>(p + (16 < Lpad ? 16 : 0));
    uint4 cur = (0 < Lpad) ? c0 : null4;
    uint4 nxt = (16 < Lpad) ? c1 : null4;
    load_e((int) (cur.x & 0xffu));
    load_carries(0);
    for (int i0 = 0; i0 < L; i0 += 16) {
      const int i2 = i0 + 32;
      const uint4 raw = *reinterpret_cast<const uint4 *>(p + (i2 < Lpad ? i2 : 0));
      // Incoming carries: M, I, D of rows i0-1 .. i0+14 (lane 0 the row
      // before the block) and the D entries of rows i0 .. i0+15.
      qM = __builtin_amdgcn_update_dpp(clM, pM, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
      qI = __builtin_amdgcn_update_dpp(clI, pI, 0x111, 0xf, 0xf, false);
      qD = __builtin_amdgcn_update_dpp(clD, pD, 0x111, 0xf, 0xf, false);
      qE = pE;
      clM = (uint32_t) __builtin_amdgcn_readlane((int) pM, 15);
      clI = (uint32_t) __builtin_amdgcn_readlane((int) pI, 15);
      clD = (uint32_t) __builtin_amdgcn_readlane((int) pD, 15);
      store_carries();
      load_carries(i0 + 16);
      uint32_t w0 = cur.x, w1 = cur.y, w2 = cur.z, w3 = cur.w;     // 16 residues, four per word
      // Rows past L up to the next multiple of four run on the padded
      // chunk's residues; they neither vote nor count.
      const int nq = min(4, (L - i0 + 3) >> 2);
      for (int qd = 0; qd < nq; ++qd) {
        const int ib = i0 + 4 * qd;
#pragma unroll
        for (int r = 0; r < 4; ++r) row((int) (((r < 3 ? w0 >> (8 * r + 8) : w1)) & 0xffu), ib + r < L);
        w0 = w1; w1 = w2; w2 = w3; w3 = nxt.x;
      }
      // Outgoing carries of the block's rows: the last row's values into
      // lane 63, then lane 64 - nr + k holds row k.
      if (!a.last) {
        sM = __builtin_amdgcn_update_dpp(Mr[Q - 1], qM, 0x101, 0xf, 0xf, false);
        sI = __builtin_amdgcn_update_dpp(Ir[Q - 1], qI, 0x101, 0xf, 0xf, false);
        sD = __builtin_amdgcn_update_dpp(Dr[Q - 1], qD, 0x101, 0xf, 0xf, false);
        sE = __builtin_amdgcn_update_dpp(dent, qE, 0x101, 0xf, 0xf, false);
        snr = 4 * nq;
        si = i0;
      }
      cur = nxt;
      nxt = (i2 < Lpad) ? raw : null4;
    }
    store_carries();
    const uint32_t xm = group_max2<G>(xacc);
    int xE = lo16(__builtin_amdgcn_readlane(xm, G - 1));
    if (!a.first) {
      const int pv = a.partial[it];
      xE = max(xE, (int) (int16_t) (pv & 0xffff));
      fired |= (pv >> 16) & 1;
    }
    if (lane == 0) {
      if (!a.last) a.partial[it] = (fired << 16) | (int) (uint16_t) xE;
      else a.out[it] = fired ? VIT_DEFER : (int) (uint16_t) max(-32768, xE + a.e_move);   // :179, C loop 0
    }
  }
}

template <int Q>
static size_t seg_lds(int R) { return band_lds<64, Q>(R); }

template <int Q>
static int cap_seg(int R, int ncu) {
  return ncu * resident_per_cu(vit_seg<Q>, 512, seg_lds<Q>(R));
}

template <int Q>
static hipError_t launch_seg(const VitSegArgs *a, int grid, hipStream_t st) {
  const size_t lds = seg_lds<Q>(a->db.R);
  hipLaunchKernelGGL(vit_seg<Q>, dim3(grid), dim3(512), lds, st, *a);
  return hipGetLastError();
}

template <int... Qs>
static hipError_t dispatch_seg(int Q, const VitSegArgs *a, int grid, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t e = hipErrorInvalidValue;
  ((Q == Qs + 1 ? (e = launch_seg<Qs + 1>(a, grid, st), 0) : 0), ...);
  return e;
}

template <int... Qs>
static int caps_seg(int Q, int R, int ncu, std::integer_sequence<int, Qs...>) {
  int r = 0;
  ((Q == Qs + 1 ? (r = cap_seg<Qs + 1>(R, ncu), 0) : 0), ...);
  return r;
}

// One column segment of Q <= 14 registers per lane (beyond, the carries'
// registers push the lane past 256 VGPRs): grid workgroups of 512 threads
// (eight waves, wave slot b * 8 + w).
extern "C" hipError_t msvk_launch_vit_seg(int Q, const VitSegArgs *a, int grid, hipStream_t st) {
  if (Q < 1 || Q > 14 || grid < 1) return hipErrorInvalidValue;
  return dispatch_seg(Q, a, grid, st, std::make_integer_sequence<int, 14>{});
}

// Workgroups of vit_seg<Q> the device holds at once for R + 1 table rows.
extern "C" int msvk_vit_seg_cap(int Q, int R, int ncu) {
  if (Q < 1 || Q > 14) return 0;
  return caps_seg(Q, R, ncu, std::make_integer_sequence<int, 14>{});
}
