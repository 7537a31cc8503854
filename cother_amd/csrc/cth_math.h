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

// cth_expf is the exponential the MAP score adjustment uses in place of
// CUDA's __expf.  Method: k = rint(x / ln2), r = x - k ln2 (ln2 in two
// parts, HI with 16 significant bits so k*HI is exact for |k| < 256),
// e^r by its Taylor polynomial to r^7 (|r| <= 0.347: truncation below
// 5e-9 relative), scaled by 2^k through the exponent bits.  Defined for
// -87 <= x <= 88; below it returns 0, above it +inf.
CTH_HD float cth_expf(float x) {
  if (!(x >= -87.0f)) return x != x ? x : 0.0f;
  if (x > 88.0f) return INFINITY;
  const float k = rintf(x * 1.44269504f);
  float r = fmaf(-k, 0.693145751953125f, x);
  r = fmaf(-k, 1.42860682e-06f, r);
  float p = fmaf(r, 1.98412698e-04f, 1.38888889e-03f);  // 1/7!, 1/6!
  p = fmaf(r, p, 8.33333333e-03f);                     // 1/5!
  p = fmaf(r, p, 4.16666667e-02f);                     // 1/4!
  p = fmaf(r, p, 1.66666667e-01f);                     // 1/3!
  p = fmaf(r, p, 0.5f);
  p = fmaf(r, p, 1.0f);
  p = fmaf(r, p, 1.0f);
  union { float f; uint32_t u; } v;
  v.u = (uint32_t)((int)k + 127) << 23;
  return p * v.f;
}

// Schraudolph's exponential (Neural Computation 11, 1999) with the
// constants of the reference's MAP kernels (cucom/myfastapproxmath.cuh:
// 12-41): the float bits int(12102203.16 x + 1064631197) with one fused
// multiply-add; 0 unless -16 < x (MAPDPfastapprox_expf,
// cumapdp/CuBatchMAPDP_fwd.cuh:204-214).
CTH_HD float cth_fastexpf(float x) {
  if (!(-16.0f < x)) return 0.0f;
  union { float f; int32_t i; } v;
  v.i = (int32_t)fmaf(12102203.16f, x, 1064631197.0f);
  return v.f;
}

// Its inverse: (bits(x) - 1064631197) * 8.2629582959339e-08; -99 unless
// 3.8e-11 < x (MAPDPfastapprox_logf, cumapdp/CuBatchMAPDP_fwd.cuh:234-240).
CTH_HD float cth_fastlogf(float x) {
  if (!(3.8e-11f < x)) return -99.0f;
  union { float f; int32_t i; } v;
  v.f = x;
  return (float)(v.i - 1064631197) * 8.2629582959339e-08f;
}
