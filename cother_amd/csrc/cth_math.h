// cth_math.h: single-precision helpers shared by host and device code.
//
// cth_logf is the logarithm the score terms use in place of CUDA's __logf
// (an NVIDIA-specific approximation that this build does not reproduce).
// Method: x = 2^e m, m in [sqrt(1/2), sqrt(2)); f = m - 1 (exact);
// s = f / (2 + f); log(1+f) = 2 atanh(s) = f - s (f - R) with
// R = sum_{k=1..5} 2 s^{2k} / (2k+1) (Taylor series of atanh, truncation
// below 1e-10 relative for |s| <= 0.1716); then e*ln2 added in two parts
// (ln2 = HI + LO, HI carrying 16 significant bits so e*HI is exact).
// Accuracy: within 1 ulp of the correctly rounded logarithm.  Every
// operation is an explicit IEEE single-precision operation, so host and
// device produce identical bits when compiled without FP contraction.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#define CTH_HD __host__ __device__ __forceinline__
#else
#define CTH_HD static inline
#endif

CTH_HD float cth_logf(float x) {
  if (!(x > 0.0f)) return x == 0.0f ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  union { float f; uint32_t u; } v;
  v.f = x;
  int e = 0;
  if (v.u < 0x00800000u) { v.f = x * 8388608.0f; e = -23; }
  e += (int)(v.u >> 23) - 127;
  v.u = (v.u & 0x007fffffu) | 0x3f800000u;
  if (v.u > 0x3fb504f3u) { v.f = v.f * 0.5f; e += 1; }
  const float f = v.f - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float r = fmaf(z, 0.18181818f, 0.22222222f);   // 2/11, 2/9
  r = fmaf(z, r, 0.28571429f);                    // 2/7
  r = fmaf(z, r, 0.4f);                           // 2/5
  r = fmaf(z, r, 0.66666669f);                    // 2/3
  r = z * r;
  const float l1p = fmaf(-s, f - r, f);
  const float fe = (float)e;
  const float lo = fmaf(fe, 1.42860682e-06f, l1p);
  return fmaf(fe, 0.693145751953125f, lo);
}
