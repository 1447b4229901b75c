// msv_host.h -- internal host helpers shared by the runtime and msv_host.cpp.
#pragma once
#include <cmath>
#include <cstdint>

// tjb_b(L) = unbiased_byteify(logf(3.0f / (L+3)))
// (impl_sse/p7_oprofile.c:921-925 with :512-519; scale_b = 3/ln2, :620).
inline int msvh_tjb_impl(int L) {
  const float scale_b = 3.0 / 0.69314718055994529;
  float sc = logf(3.0f / (float) (L + 3));
  sc = -1.0f * roundf(scale_b * sc);
  return (sc > 255.) ? 255 : (uint8_t) sc;
}

// nullsc(L) after p7_bg_SetLength(bg, L) (p7_bg.c:191-198) and
// p7_bg_NullOne (p7_bg.c:359-363).
inline float msvh_nullsc_impl(int L) {
  const float p1 = (float) L / (float) (L + 1);
  return (float) L * log(p1) + log(1. - p1);
}

// The same two functions through the C ABI names used by the runtime.
extern "C" int msvh_tjb(int L);
extern "C" float msvh_nullsc(int L);
