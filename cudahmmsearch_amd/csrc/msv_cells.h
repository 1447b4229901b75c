// msv_cells.h -- device helpers shared by the MSV kernels: the packed f16
// cell operations, the one-slot diagonal shift, table-row loads from LDS,
// the decision rules of the J-free pass and the deferred-pair queues.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>
#include "msv_device.h"

typedef _Float16 h2v __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

namespace {

__device__ __forceinline__ h2v as_h2(uint32_t v) { return __builtin_bit_cast(h2v, v); }
__device__ __forceinline__ uint32_t h2_u32(h2v v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ u16x2 as_u2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t u2_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// One DP cell pair: clamp(W' + e) to [0, 1] -> v_pk_add_f16 ... clamp.
__device__ __forceinline__ uint32_t cell(uint32_t w, uint32_t e) {
  const h2v zero = {(_Float16)0.0f, (_Float16)0.0f};
  const h2v one = {(_Float16)1.0f, (_Float16)1.0f};
  h2v s = as_h2(w) + as_h2(e);
  s = __builtin_elementwise_min(__builtin_elementwise_max(s, zero), one);
  return h2_u32(s);
}
// Running maximum over cells -> v_pk_maximum3_f16.
__device__ __forceinline__ uint32_t max3h(uint32_t a, uint32_t b, uint32_t c) {
  return h2_u32(__builtin_elementwise_maximum(__builtin_elementwise_maximum(as_h2(a), as_h2(b)), as_h2(c)));
}
__device__ __forceinline__ uint32_t max2h(uint32_t a, uint32_t b) {
  return h2_u32(__builtin_elementwise_maximum(as_h2(a), as_h2(b)));
}

// Value entering slot 0 of register 0 from slot 2G-1 of the last register,
// per lane group: slot s <- slot s-1, group-first slot <- 0 (the floor).
template <int G>
__device__ __forceinline__ uint32_t shift_slot(uint32_t last, int lane) {
  uint32_t prev;
  if constexpr (G == 64) {
    prev = __builtin_amdgcn_update_dpp(0u, last, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
  } else if constexpr (G == 16) {
    prev = __builtin_amdgcn_update_dpp(0u, last, 0x111 /* row_shr:1 */, 0xf, 0xf, true);
  } else if constexpr (G == 8) {
    prev = __builtin_amdgcn_update_dpp(0u, last, 0x111 /* row_shr:1 */, 0xf, 0xf, true);
    prev = (lane & 7) ? prev : 0u;       // the second group of each DPP row starts at lane 8
  } else {
    prev = __builtin_amdgcn_update_dpp(0u, last, 0x138 /* wave_shr:1 */, 0xf, 0xf, true);
    prev = (lane == 32) ? 0u : prev;
  }
  return __builtin_amdgcn_alignbit(last, prev, 16);
}

// Wave-uniform load through the scalar cache (read-only data only).  The
// host pass of the compiler only parses device code; it gets a plain load.
template <class T>
__device__ __forceinline__ T sload(const T *p) {
#if __HIP_DEVICE_COMPILE__
  return *(const __attribute__((address_space(4))) T *) p;
#else
  return *p;
#endif
}

__device__ __forceinline__ uint32_t byte_of(const uint4 &c, int rr) {
  const uint32_t w = (rr < 4) ? c.x : (rr < 8) ? c.y : (rr < 12) ? c.z : c.w;
  return (w >> (8 * (rr & 3))) & 0xffu;
}

// v_writelane_b32: lane `lane` of dst takes the wave-uniform value val.
template <int lane>
__device__ __forceinline__ uint32_t write_lane(uint32_t dst, uint32_t val) {
  asm("v_writelane_b32 %0, %1, %2" : "+v"(dst) : "s"(val), "i"(lane));
  return dst;
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [I0, N).
template <int I0, int N, class F>
__device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I0 < N) {
    f(std::integral_constant<int, I0>{});
    static_for<I0 + 1, N>(f);
  }
}

// Ends a DP row: the row's cells are materialized here and nothing of the
// next row is computed before (stops the compiler from interleaving unrolled
// rows along diagonals, which keeps sixteen rows of cells live at once).
template <int Q>
__device__ __forceinline__ void row_fence(uint32_t (&r)[Q]) {
#pragma unroll
  for (int j = 0; j < Q; ++j) asm volatile("" : "+v"(r[j]));
}

// The Q registers' emissions of residue row x for this lane: ceil(Q/4)
// 16-byte table entries, the last one read only as wide as it is used
// (the table is the co-limiting LDS stream, so unused bytes are not read).
// Entries e[Q..] are left unset.
// TW is the table's width in lanes: 64, or G when the lane groups of a wave
// share one copy (entry of lane l at l % G).
template <int Q, int TW = 64>
__device__ __forceinline__ void load_row(const uint4 *tl, uint32_t x, uint32_t (&e)[((Q + 3) / 4) * 4]) {
  constexpr int QB = (Q + 3) / 4, TAIL = Q - 4 * (QB - 1);
  const uint4 *tp = tl + x * (QB * TW);
#pragma unroll
  for (int jb = 0; jb < QB - 1; ++jb) {
    const uint4 v = tp[jb * TW];
    e[4 * jb + 0] = v.x; e[4 * jb + 1] = v.y; e[4 * jb + 2] = v.z; e[4 * jb + 3] = v.w;
  }
  const uint4 *lp = tp + (QB - 1) * TW;
  constexpr int b = 4 * (QB - 1);
  if constexpr (TAIL == 4) {
    const uint4 v = *lp;
    e[b] = v.x; e[b + 1] = v.y; e[b + 2] = v.z; e[b + 3] = v.w;
  } else if constexpr (TAIL == 3) {
    const uint2 v = *reinterpret_cast<const uint2 *>(lp);
    e[b] = v.x; e[b + 1] = v.y;
    e[b + 2] = reinterpret_cast<const uint32_t *>(lp)[2];
  } else if constexpr (TAIL == 2) {
    const uint2 v = *reinterpret_cast<const uint2 *>(lp);
    e[b] = v.x; e[b + 1] = v.y;
  } else {
    e[b] = *reinterpret_cast<const uint32_t *>(lp);
  }
}


// The decision the reference reaches from the J-free maximum, returning the
// result word.  Rp is max(W - B0) over all cells, in units (0..2048).
//   SSV regime (tjb+tbm+tec+bias < 127), ssvfilter.c:1370-1410:
//     overflow -> eslERANGE; J possible (xE - tec > base) -> byte MSV;
//     else xJ = max(B0, cells) - tec.
//   Guard regime: the byte MSV itself; while J never fires its B is the
//   constant B0 of msvfilter.c:124, so its cells are the same J-free cells.
__device__ __forceinline__ uint32_t decide(int Rp, int tjb, const ProfDev &pd) {
  const int tjbm = (tjb + pd.tbm) & 255;             // msvfilter.c:120
  const int B0 = max(0, pd.base - tjbm);             // msvfilter.c:124
  const int R = Rp + B0;
  const bool guard = (tjb + pd.tbm + pd.tec + pd.bias) >= 127;
  if (pd.slow_all) return RES_SLOW << 8;
  if (R >= 255 - pd.bias) return RES_ERANGE << 8;
  if (R - pd.tec > pd.base) return (RES_SLOW << 8) | 1u;  // only J is in doubt
  if (guard) {
    if (Rp == 0) return RES_SLOW << 8;                // exact sub-B0 maximum unknown
    return (uint32_t) max(0, R - pd.tec);
  }
  return (uint32_t) (R - pd.tec);
}

// Queue a deferred pair: J-only doubts go to the profile's J-aware queue when
// it has one, everything else to the byte-MSV queue.
__device__ __forceinline__ void defer(uint32_t w, int prof_idx, int64_t s, SlowItem *slow, int32_t *slow_count,
                                      int32_t slow_cap, SlowItem *jslow, int32_t *jslow_count, int32_t jslow_cap) {
  if ((w & 1u) && jslow) {
    const int k = atomicAdd(jslow_count, 1);
    if (k < jslow_cap) { jslow[k] = SlowItem{prof_idx, (int32_t) s}; return; }
  }
  const int k = atomicAdd(slow_count, 1);
  if (k < slow_cap) slow[k] = SlowItem{prof_idx, (int32_t) s};
}

// Wave maximum of non-negative integers by the same DPP steps (the zeros
// that bound_ctrl shifts in never win); every lane gets it.
__device__ __forceinline__ int wave_max_nonneg(int v) {
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false));
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false));
  return __builtin_amdgcn_readlane(v, 63);
}

__device__ __forceinline__ float wave_max_f32(float v) {
  // Row-wise DPP reduction, then across rows; lane 63 ends with the maximum.
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true)));
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, true)));
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, true)));
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, true)));
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xa, 0xf, false)));
  v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xc, 0xf, false)));
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

}  // namespace
