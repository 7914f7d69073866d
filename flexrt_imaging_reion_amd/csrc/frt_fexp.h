// frt_fexp.h -- FP32 exp(-d) and expm1(-d) from one range reduction.
//
// Used by the mixed-precision deposit, where each segment and bin needs the
// transmission exp(-dtau) and the absorbed fraction 1 - exp(-dtau) of two
// optical depths.  Evaluating both from one reduction and one polynomial
// replaces separate expf/expm1f calls.  Host-callable so the CPU tests can
// check it against double-precision exp/expm1.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef __HIPCC__
#define FRT_FEXP_HD __host__ __device__
#else
#define FRT_FEXP_HD
#endif

// 2^k as a float for -126 <= k <= 127 (exact).
FRT_FEXP_HD inline float frt_pow2f(int k) {
  const uint32_t bits = (uint32_t)(k + 127) << 23;
  float f;
  memcpy(&f, &bits, 4);
  return f;
}

// For d >= 0: e = exp(-d) and m = exp(-d) - 1.
//   -d = k ln2 + r, |r| <= ln2/2, with ln2 split Cody-Waite style
//   (ln2_hi has 16 significant bits, so k * ln2_hi is exact for |k| <= 127);
//   q = expm1(r) by its Taylor series through r^7 (truncation below 2e-8
//   relative for |r| <= 0.35);
//   exp(-d) = 2^k (1 + q),  expm1(-d) = 2^k q + (2^k - 1)  (one fma).
// Below d = 1/32 (most segments of an ionized region) the reduction is
// skipped: 1 - exp(-d) = d (1 - d/2 + d^2/6 - d^3/24 + d^4/120) with the
// next term below 1.3e-9 relative, and exp(-d) = 1 - that.
// d = 0 gives e = 1, m = 0 exactly.  Above d = 87 exp(-d) leaves the normal
// range: e = 0 and m = -1 are returned.  NaN propagates.
FRT_FEXP_HD inline void frt_exp_expm1_neg(float d, float& e, float& m) {
  if (d < 0.03125f) {
    float q = fmaf(d, 1.f / 120.f, -1.f / 24.f);
    q = fmaf(q, d, 1.f / 6.f);
    q = fmaf(q, d, -0.5f);
    q = fmaf(q, d, 1.f);
    m = fmaf(-d, q, 0.f);  // +0 at d = 0
    e = 1.f + m;
    return;
  }
  if (d > 87.f) {
    e = 0.f;
    m = -1.f;
    return;
  }
  const float x = -d;
  const float kf = rintf(x * 1.44269504088896341f);
  float r = fmaf(kf, -0.693145751953125f, x);
  r = fmaf(kf, -1.42860676533018714e-06f, r);
  float p = 1.f / 5040.f;
  p = fmaf(p, r, 1.f / 720.f);
  p = fmaf(p, r, 1.f / 120.f);
  p = fmaf(p, r, 1.f / 24.f);
  p = fmaf(p, r, 1.f / 6.f);
  p = fmaf(p, r, 0.5f);
  const float q = fmaf(p, r * r, r);
  const float s = frt_pow2f((int)kf);
  e = s * (1.f + q);
  m = fmaf(s, q, s - 1.f);
}
