// vit_kernels.hip -- gfx950 kernel for HMMER3's ViterbiFilter, the filter
// stage after MSV in p7_Pipeline (impl_sse/vitfilter.c:84-250), bit-exact.
//
// The reference runs Farrar's striped int16 recurrence over 8 words per SSE
// vector.  Here W wavefronts (one when M <= 1024) own a target and stripe its
// model over 128*W int16 slots, two per VGPR: register j of lane c holds model
// positions k = j + s*Q + 1 for its slots s = 2c, 2c+1.  Every cell does the
// reference's operations in the reference's order with packed saturating
// int16 arithmetic (v_pk_add_i16 clamp = _mm_adds_epi16, v_pk_max_i16 =
// _mm_max_epi16), so values do not depend on the striping.  Per row:
//   - the diagonal predecessors of register 0 come from the last register one
//     slot down: DPP wave_shr:1 + v_alignbit, and across waves through LDS;
//   - xE and Dmax (vitfilter.c:176, :198) are one packed DPP max reduction;
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

namespace {
template <int G> __device__ __forceinline__ uint32_t shift_grp(uint32_t last, int lane);
template <int G, int Q>
__device__ __forceinline__ void close_d(uint32_t (&Dr)[Q], const uint32_t (&tDD)[Q], bool mine, const uint4 *tsc,
                                        int c, int lane);
}  // namespace

template <int Q, int W>
__global__ void __launch_bounds__(W == 1 ? 256 : 64 * W) vit_kernel(VitArgs a) {
  extern __shared__ uint32_t etab_lds[];               // one wave per target: the emission table
  __shared__ uint32_t red[16], dlast[16], eM[16], eI[16], eD[16];
  __shared__ int scC[16], scT[16];
  constexpr int C = 64 * W;                            // lanes per target
  __shared__ uint4 tsc[C];                             // the lane scan's T terms per lane
  __shared__ uint32_t tse[W > 1 ? C : 1];              // and its cross-wave terms
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int w = (W == 1) ? 0 : wave;
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
  if (W > 1 || threadIdx.x < 64) {
    tsc[c] = make_uint4(negsat2(Tlo, Thi), negsat2(in_row_t<1>(Ts0, lane), in_row_t<2>(Ts1, lane)),
                        negsat2(in_row_t<4>(Ts2, lane), in_row_t<8>(Ts3, lane)), negsat2(Ts4, Ts5));
    if constexpr (W > 1) tse[c] = negsat2(Ts6, Tex);
  }
  __syncthreads();

  const uint32_t *et = a.etab;
  if (W == 1 && a.etab_in_lds) {
    const int ntab = (a.db.R + 1) * Q * 64;
    for (int i = threadIdx.x; i < ntab; i += blockDim.x) etab_lds[i] = a.etab[i];
    __syncthreads();
    et = etab_lds;
  }

  const int nw = (W == 1) ? (int) (blockDim.x >> 6) : 1;
  const int first = (W == 1) ? (int) blockIdx.x * nw + wave : (int) blockIdx.x;
  const int stride = (W == 1) ? (int) gridDim.x * nw : (int) gridDim.x;
  const int xN = a.base_w;                             // vitfilter.c:116; N loop cost is 0

  for (int it = first; it < a.n; it += stride) {
    const int64_t s = a.items[it];
    const int L = a.db.len[s];
    const uint8_t *p = a.db.res + a.db.off[s];
    const int move = a.move[it];
    uint32_t Mr[Q], Ir[Q], Dr[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) Mr[j] = Ir[j] = Dr[j] = NEG2;        // :115
    int xB = (int16_t) (xN + move);                                  // :117
    int xJ = -32768, xC = -32768, status = 0;
    int xacc = -32768;                  // one wave: this lane's largest row maximum (xC at the end)
    // one wave: the row votes' thresholds (as in vit_group)
    int tE = min(max(xJ, xN) - a.e_loop, 32766), tD = xB - a.ddbound_w;
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
        if constexpr (W == 1) {
          // This lane's share of the row's xE and Dmax (:176, :198), voted
          // on instead of reduced: the exact special-state update runs only
          // when some lane could overflow or move B.
          const uint32_t hv = halves_max(xEv, Dmv);
          const int xe = lo16(hv), dm = hi16(hv);
          xacc = max(xacc, xe);
          if (__builtin_amdgcn_ballot_w64(xe > tE)) {
            const int xE = lo16(wave_max2(hv));
            if (xE >= 32767) { status = 16; break; }                   // :177
            xJ = (int16_t) max(xJ, xE + a.e_loop);                     // :180 (J loop 0)
            const int xBn = (int16_t) max(xJ + move, xN + move);       // :181
            if (xBn != xB) {
#pragma unroll
              for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xBn, xBn), tBM[j]);
            }
            xB = xBn;
            tE = min(max(xJ, xN) - a.e_loop, 32766);
            tD = xB - a.ddbound_w;
          }
          Dr[0] = shift_in(dcv, NEG2, lane);                           // :230-231 / :203-204
          if (__builtin_amdgcn_ballot_w64(dm > tD)) close_d<64, Q>(Dr, tDD, true, tsc, c, lane);   // :199-227
          continue;
        }
        uint32_t r = wave_max2(halves_max(xEv, Dmv));
        uint32_t dedge = NEG2;
        if constexpr (W > 1) {
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
          int cin_lo = ce;             // one wave: nothing enters its first slot
          if constexpr (W > 1) {
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
        if constexpr (W > 1) {
          if (lane == 63) { eM[w] = Mr[Q - 1]; eI[w] = Ir[Q - 1]; eD[w] = Dr[Q - 1]; }
          __syncthreads();
          if (w > 0) { edM = eM[w - 1]; edI = eI[w - 1]; edD = eD[w - 1]; }
        }
      }
    }
    if constexpr (W == 1) {
      // xC = max over the rows of xE + e_move (:179, C loop 0)
      xC = max(-32768, lo16(wave_max2(pack16(xacc, xacc))) + a.e_move);
    }
    // Several waves: every wave has read this item's last exchange (red[]
    // on a row that overflowed) before any wave's next item writes it.  The
    // item, its length and the overflow break are the same for all waves
    // of the workgroup, so all of them reach this barrier once per item.
    if constexpr (W > 1) __syncthreads();
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

// Grids: one wave per target (or per NG targets) with a static stride, as
// many workgroups as are needed or as fit the device at once, whichever is
// fewer; multi-wave targets keep one workgroup per target up to 8 waves per
// CU.
template <int Q, int W>
static int grid_vit(const VitArgs *a, int ncu) {
  const int block = (W == 1) ? 256 : 64 * W;
  const size_t lds = (W == 1 && a->etab_in_lds) ? (size_t) (a->db.R + 1) * Q * 64 * 4 : 0;
  if (W == 1) {
    const int cap = ncu * resident_per_cu(vit_kernel<Q, W>, block, lds);
    return std::max(1, std::min((a->n + 3) / 4, cap));
  }
  return std::max(1, std::min(a->n, ncu * std::max(1, 8 / W)));
}

template <int Q, int W>
static hipError_t launch_vit(const VitArgs *a, int ncu, hipStream_t st) {
  const int block = (W == 1) ? 256 : 64 * W;
  const size_t lds = (W == 1 && a->etab_in_lds) ? (size_t) (a->db.R + 1) * Q * 64 * 4 : 0;
  const int grid = grid_vit<Q, W>(a, ncu);
  hipLaunchKernelGGL((vit_kernel<Q, W>), dim3(grid), dim3(block), lds, st, *a);
  return hipGetLastError();
}

template <int W, int... Qs>
static hipError_t dispatch_q(int Q, const VitArgs *a, int ncu, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t e = hipErrorInvalidValue;
  ((Q == Qs + 1 ? (e = launch_vit<Qs + 1, W>(a, ncu, st), 0) : 0), ...);
  return e;
}

// Q <= 16 registers per lane and W = 1, 2, 4 or 8 waves per target (M up to
// 16384): the transitions of 16 registers still fit the 256 VGPRs a lane
// has when a workgroup holds at most two waves per SIMD.
extern "C" hipError_t msvk_launch_vit(int Q, int W, const VitArgs *a, int ncu, hipStream_t st) {
  if (Q < 1 || Q > 16 || a->n < 1) return hipErrorInvalidValue;
  auto qs = std::make_integer_sequence<int, 16>{};
  switch (W) {
    case 1: return dispatch_q<1>(Q, a, ncu, st, qs);
    case 2: return dispatch_q<2>(Q, a, ncu, st, qs);
    case 4: return dispatch_q<4>(Q, a, ncu, st, qs);
    case 8: return dispatch_q<8>(Q, a, ncu, st, qs);
    default: return hipErrorInvalidValue;
  }
}
// ---------------------------------------------------------------------------
// Short models (M <= 512): NG = 64/G targets per wave, G = 16 or 32 lanes
// each (register j of lane c of a group holds k = j + s*Q + 1 for its slots
// s = 2c, 2c+1), so the per-row fixed work -- the reductions, the shifts,
// the special states, the delete-closure scan -- is shared by NG targets.
// Cells are the same operations as vit_kernel in the same order.  The
// special states of each group are wave-uniform values (one set per group);
// a group whose target has ended (its rows are done, or it overflowed)
// keeps computing padding rows but no longer updates its states, and its
// closure is never applied.  The emission table (G lanes wide) is staged in
// LDS; items are taken NG at a time by a static stride, longest first.
// ---------------------------------------------------------------------------
namespace {

// Slot s <- slot s-1 inside each G-lane group; the group's first slot takes
// minus infinity.
template <int G>
__device__ __forceinline__ uint32_t shift_grp(uint32_t last, int lane) {
  uint32_t prev;
  if constexpr (G == 16) {
    prev = __builtin_amdgcn_update_dpp(NEG2, last, 0x111 /* row_shr:1 */, 0xf, 0xf, false);
  } else {
    prev = __builtin_amdgcn_mov_dpp(last, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
    prev = ((lane & (G - 1)) == 0) ? NEG2 : prev;
  }
  return __builtin_amdgcn_alignbit(last, prev, 16);
}

// Ballot bits of one G-lane group.
template <int G>
constexpr uint64_t GMASK = (G == 64) ? ~0ull : ((1ull << G) - 1);

// The delete closure the reference's lazy-F passes compute (vitfilter.c:
// 203-227), for the lanes whose group runs them (mine).  Dr[0] already
// holds the D entering each slot's first position.  One pass along each
// slot (the reference's first pass); then, if the value leaving some slot
// raises the next slot's first D, the carries across slots follow from
// the max-plus scan over the lanes (tt: the lane's T terms, see tsc) and a
// second pass applies them.  With slots of only Q positions the carries
// cross slot boundaries often, so repeating whole passes (the reference
// stops at the first position no slot raises) would take many passes.
template <int G, int Q>
__device__ __forceinline__ void close_d(uint32_t (&Dr)[Q], const uint32_t (&tDD)[Q], bool mine, const uint4 *tsc,
                                        int c, int lane) {
  if (mine) {
#pragma unroll
    for (int j = 1; j < Q; ++j) Dr[j] = maxs(Dr[j], adds(Dr[j - 1], tDD[j - 1]));
  }
  const uint32_t outl = adds(Dr[Q - 1], tDD[Q - 1]);
  const uint32_t d0 = maxs(Dr[0], shift_grp<G>(outl, lane));
  if (__builtin_amdgcn_ballot_w64(mine && d0 != Dr[0]) == 0) return;
  const uint4 tt = tsc[c];
  const int clo = lo16(outl);
  const uint32_t tsl = tt.x, ts01 = tt.y, ts23 = tt.z, ts45 = tt.w;
  // the lane's two slots as one map x -> max(c, x + T): low slot, then high
  int cs = max(hi16(outl), clo - (int) (tsl >> 16));          // clo through the high slot
  // inclusive scan over the lanes of each group (lane l: its group's lanes
  // up to l composed, in order): row_shr 1/2/4/8 inside rows of 16 lanes,
  // then row_bcast 15 (and 31) across rows; lanes without a source take the
  // identity map (-32768, 0), which leaves them unchanged
  cs = c_step_row<0x111, 0>(cs, ts01);
  cs = c_step_row<0x112, 1>(cs, ts01);
  cs = c_step_row<0x114, 0>(cs, ts23);
  cs = c_step_row<0x118, 1>(cs, ts23);
  if constexpr (G >= 32) cs = c_step<0x142, 0xa, 0>(cs, ts45);
  if constexpr (G == 64) cs = c_step<0x143, 0xc, 1>(cs, ts45);
  // exclusive: the previous lane's prefix; a group's first lane the identity
  int ce;
  if constexpr (G == 16) {
    ce = __builtin_amdgcn_update_dpp(-32768, cs, 0x111, 0xf, 0xf, false);
  } else {
    ce = __builtin_amdgcn_update_dpp(-32768, cs, 0x138, 0xf, 0xf, false);
    if (G == 32 && (lane & 31) == 0) ce = -32768;
  }
  const int cin_hi = max(clo, ce - (int) (tsl & 0xffffu));
  if (mine) {
    uint32_t y = pack16(ce, cin_hi);
#pragma unroll
    for (int j = 0; j < Q; ++j) { Dr[j] = maxs(Dr[j], y); y = adds(y, tDD[j]); }
  }
}

// Packed int16 max over each G-lane group; the group's last lane holds it.
template <int G>
__device__ __forceinline__ uint32_t group_max2(uint32_t v) {
  v = row_max2(v);
  if constexpr (G == 32) v = maxs(v, __builtin_amdgcn_update_dpp(NEG2, v, 0x142 /* row_bcast:15 */, 0xa, 0xf, false));
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

}  // namespace

template <int G, int Q>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Q <= 7 ? 4 : 1, 8))) vit_group(VitArgs a) {
  extern __shared__ uint32_t etab_lds[];
  constexpr int NG = 64 / G;
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

  // tBM stays in the table: it is read only on the rows where a group's xB
  // changes (a few per target), and leaving it out of the registers keeps
  // four waves per SIMD at Q = 7
  uint32_t tMM[Q], tIM[Q], tDM[Q], tMD[Q], tMI[Q], tII[Q], tDD[Q];
#pragma unroll
  for (int j = 0; j < Q; ++j) {
    tMM[j] = a.ttab[(1 * Q + j) * G + c];
    tIM[j] = a.ttab[(2 * Q + j) * G + c];
    tDM[j] = a.ttab[(3 * Q + j) * G + c];
    tMD[j] = a.ttab[(4 * Q + j) * G + c];
    tMI[j] = a.ttab[(5 * Q + j) * G + c];
    tII[j] = a.ttab[(6 * Q + j) * G + c];
    tDD[j] = a.ttab[(7 * Q + j) * G + c];
  }
  int Tlo = 0, Thi = 0;                                // tDD summed over each slot
#pragma unroll
  for (int j = 0; j < Q; ++j) { Tlo += lo16(tDD[j]); Thi += hi16(tDD[j]); }
  // the lane scan's T terms, once: the lane's two slots and before each step
  const int Ts0 = Tlo + Thi;
  const int Ts1 = t_next<0x111, 0xf>(Ts0);
  const int Ts2 = t_next<0x112, 0xf>(Ts1);
  const int Ts3 = t_next<0x114, 0xf>(Ts2);
  const int Ts4 = (G == 32) ? t_next<0x118, 0xf>(Ts3) : 0;
  // kept in LDS, read on the rows that run the closure
  __shared__ uint4 tsc[G];
  if (threadIdx.x < G)
    tsc[c] = make_uint4(negsat2(Tlo, Thi), negsat2(in_row_t<1>(Ts0, lane), in_row_t<2>(Ts1, lane)),
                        negsat2(in_row_t<4>(Ts2, lane), in_row_t<8>(Ts3, lane)), negsat(Ts4));

  const int ntab = (a.db.R + 1) * Q * G;
  for (int i = threadIdx.x; i < ntab; i += blockDim.x) etab_lds[i] = a.etab[i];
  __syncthreads();

  const uint32_t nb = (uint32_t) a.db.R * 0x01010101u;
  const uint4 null4 = make_uint4(nb, nb, nb, nb);
  const int xN = a.base_w;                             // vitfilter.c:116; N loop cost is 0
  const int nw = blockDim.x >> 6;
  const int nwi = (a.n + NG - 1) / NG;

  for (int wi = blockIdx.x * nw + wave; wi < nwi; wi += gridDim.x * nw) {
    const int it = wi * NG + g;
    const bool valid = it < a.n;
    const int64_t s = valid ? a.items[it] : 0;
    const int L = valid ? a.db.len[s] : 0;
    const int Lpad = (L + 15) & ~15;
    const uint8_t *p = a.db.res + (valid ? a.db.off[s] : 0);
    const int mvl = valid ? (int) a.move[it] : 0;

    int mv[NG], xB[NG], xJ[NG], st[NG];
    int Lmax = 0;
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      Lmax = max(Lmax, __builtin_amdgcn_readlane(L, gg * G));
      mv[gg] = __builtin_amdgcn_readlane(mvl, gg * G);
      xB[gg] = (int16_t) (xN + mv[gg]);                              // :117
      xJ[gg] = -32768; st[gg] = 0;
    }
    uint32_t Mr[Q], Ir[Q], Dr[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) Mr[j] = Ir[j] = Dr[j] = NEG2;        // :115
    // Per group: xB + tBM per register (recomputed only when xB changes),
    // and the two thresholds of the row votes (special_thresholds).
    uint32_t bm[Q];
    int tE, tD;
    auto refresh = [&]() {
      int te[NG], td[NG];
#pragma unroll
      for (int gg = 0; gg < NG; ++gg) {
        te[gg] = st[gg] ? 0x7fffffff : min(max(xJ[gg], xN) - a.e_loop, 32766);
        td[gg] = xB[gg] - a.ddbound_w;
      }
      tE = of_group<NG>(te, ge);
      tD = of_group<NG>(td, ge);
      const int xb = of_group<NG>(xB, ge);
#pragma unroll
      for (int j = 0; j < Q; ++j) bm[j] = adds(pack16(xb, xb), a.ttab[j * G + c]);
    };
    refresh();
    int xacc = -32768;                  // this lane's largest row maximum (xC at the end)

    uint4 cnext = *reinterpret_cast<const uint4 *>(p);
    for (int i0 = 0; i0 < Lmax; i0 += 16) {
      const uint4 ch = (i0 < Lpad) ? cnext : null4;
      const int i1 = i0 + 16;
      cnext = *reinterpret_cast<const uint4 *>(p + (i1 < Lpad ? i1 : 0));
      uint32_t w0 = ch.x, w1 = ch.y, w2 = ch.z, w3 = ch.w;   // 16 residues, consumed from the bottom
      const int nr = min(16, Lmax - i0);
      for (int rr = 0; rr < nr; ++rr) {
        const int i = i0 + rr;
        const int x = (int) (w0 & 0xffu);
        w0 = __builtin_amdgcn_alignbit(w1, w0, 8);
        w1 = __builtin_amdgcn_alignbit(w2, w1, 8);
        w2 = __builtin_amdgcn_alignbit(w3, w2, 8);
        w3 >>= 8;
        const uint32_t *ep = etab_lds + x * Q * G + c;
        uint32_t e[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) e[j] = ep[j * G];
        uint32_t mpv = shift_grp<G>(Mr[Q - 1], lane);               // :139-141
        uint32_t ipv = shift_grp<G>(Ir[Q - 1], lane);
        uint32_t dpv = shift_grp<G>(Dr[Q - 1], lane);
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
        // This lane's share of the row's xE and Dmax (:176, :198), voted on
        // instead of reduced: the exact special-state update runs only when
        // some lane of a group could overflow or move the group's B.
        const uint32_t hv = halves_max(xEv, Dmv);
        const int xe = lo16(hv), dm = hi16(hv);
        const bool alive = i < L;                                    // the lane's target has this row
        xacc = max(xacc, alive ? xe : -32768);
        const uint64_t mE = __builtin_amdgcn_ballot_w64(alive && xe > tE);
        if (mE) {
          const uint32_t r = group_max2<G>(hv);
#pragma unroll
          for (int gg = 0; gg < NG; ++gg) {
            if (((mE >> (gg * G)) & GMASK<G>) == 0) continue;
            const int xE = lo16(__builtin_amdgcn_readlane(r, gg * G + G - 1));
            if (xE >= 32767) { st[gg] = 16; continue; }                  // :177
            xJ[gg] = (int16_t) max(xJ[gg], xE + a.e_loop);               // :180 (J loop 0)
            xB[gg] = (int16_t) max(xJ[gg] + mv[gg], xN + mv[gg]);        // :181
          }
          refresh();
        }
        Dr[0] = shift_grp<G>(dcv, lane);                             // :230-231 / :203-204
        const uint64_t mD = __builtin_amdgcn_ballot_w64(alive && dm > tD);   // :199, per group
        if (mD) {
          int cond[NG];
#pragma unroll
          for (int gg = 0; gg < NG; ++gg) cond[gg] = (((mD >> (gg * G)) & GMASK<G>) != 0 && st[gg] == 0) ? 1 : 0;
          const bool mine = of_group<NG>(cond, ge) != 0;
          close_d<G, Q>(Dr, tDD, mine, tsc, c, lane);                 // :203-227
        }
      }
    }
    // xC = max over the target's rows of xE + e_move (:179, C loop 0): the
    // group's largest row maximum plus e_move, floored at xC's start value
    const uint32_t xm = group_max2<G>(pack16(xacc, xacc));
#pragma unroll
    for (int gg = 0; gg < NG; ++gg) {
      const int xC = max(-32768, lo16(__builtin_amdgcn_readlane(xm, gg * G + G - 1)) + a.e_move);
      if (lane == gg * G && valid) a.out[it] = (st[gg] << 16) | (int) (uint16_t) xC;
    }
  }
}

template <int G, int Q>
static int grid_group(const VitArgs *a, int ncu) {
  constexpr int NG = 64 / G;
  const size_t lds = (size_t) (a->db.R + 1) * Q * G * 4;
  const int waves = (a->n + NG - 1) / NG;
  const int cap = ncu * resident_per_cu(vit_group<G, Q>, 256, lds);
  return std::max(1, std::min((waves + 3) / 4, cap));
}

template <int G, int Q>
static hipError_t launch_group(const VitArgs *a, int ncu, hipStream_t st) {
  const size_t lds = (size_t) (a->db.R + 1) * Q * G * 4;
  const int grid = grid_group<G, Q>(a, ncu);
  hipLaunchKernelGGL((vit_group<G, Q>), dim3(grid), dim3(256), lds, st, *a);
  return hipGetLastError();
}

template <int G, int... Qs>
static hipError_t dispatch_group(int Q, const VitArgs *a, int ncu, hipStream_t st, std::integer_sequence<int, Qs...>) {
  hipError_t e = hipErrorInvalidValue;
  ((Q == Qs + 1 ? (e = launch_group<G, Qs + 1>(a, ncu, st), 0) : 0), ...);
  return e;
}

extern "C" hipError_t msvk_launch_vit_group(int G, int Q, const VitArgs *a, int ncu, hipStream_t st) {
  if (Q < 1 || Q > 8 || a->n < 1) return hipErrorInvalidValue;
  auto qs = std::make_integer_sequence<int, 8>{};
  switch (G) {
    case 16: return dispatch_group<16>(Q, a, ncu, st, qs);
    case 32: return dispatch_group<32>(Q, a, ncu, st, qs);
    default: return hipErrorInvalidValue;
  }
}

// The static schedule of a launch: item position q (in a.items order) goes
// to slot (q / per) % slots, where per targets share a wave (64/G for the
// grouped shapes, else 1) and slots is the number of waves (workgroups for
// multi-wave targets) striding over the items.  The host orders the items
// so that every slot gets about the same number of rows.
template <int... Qs>
static int slots_vit(int Q, int W, const VitArgs *a, int ncu, std::integer_sequence<int, Qs...>) {
  int r = 0;
  ((Q == Qs + 1 ? (r = (W == 1) ? grid_vit<Qs + 1, 1>(a, ncu) * 4
                    : (W == 2) ? grid_vit<Qs + 1, 2>(a, ncu)
                    : (W == 4) ? grid_vit<Qs + 1, 4>(a, ncu) : grid_vit<Qs + 1, 8>(a, ncu), 0) : 0), ...);
  return r;
}
template <int G, int... Qs>
static int slots_group(int Q, const VitArgs *a, int ncu, std::integer_sequence<int, Qs...>) {
  int r = 0;
  ((Q == Qs + 1 ? (r = grid_group<G, Qs + 1>(a, ncu) * 4, 0) : 0), ...);
  return r;
}

extern "C" int msvk_vit_slots(int G, int Q, int W, const VitArgs *a, int ncu, int *per) {
  if (a->n < 1) return 0;
  if (G < 64) {
    *per = 64 / G;
    if (Q < 1 || Q > 8) return 0;
    return G == 16 ? slots_group<16>(Q, a, ncu, std::make_integer_sequence<int, 8>{})
                   : slots_group<32>(Q, a, ncu, std::make_integer_sequence<int, 8>{});
  }
  *per = 1;
  if (Q < 1 || Q > 16) return 0;
  return slots_vit(Q, W, a, ncu, std::make_integer_sequence<int, 16>{});
}
