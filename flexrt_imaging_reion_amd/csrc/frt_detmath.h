// frt_detmath.h -- deterministic transcendental functions for host and device.
//
// Every transcendental call on the radiative-transfer path (exp, log, pow,
// sin, cos, atan2) is evaluated through these routines, on the CPU and on the
// GPU alike.  They use only IEEE-754 +, -, *, /, sqrt, rint and explicit
// fused multiply-add, each of which is correctly rounded on x86-64 and on
// gfx950, so the same input gives the same bits on both sides.  That is what
// lets the GPU path reproduce the CPU restatement bit for bit, and with it the
// threshold decisions (ray survival, split, merge, HEALPix bins) that decide
// the ray counts.
//
// Accuracy (checked against glibc in tests/test_detmath.py): exp, log, sin,
// cos, atan2 within 2 ulp; pow within 2 ulp + |y ln x| * 2^-52 relative.
//
// Build rule: compile every translation unit that includes this header with
// -ffp-contract=off, so the compiler does not fuse a*b+c on its own.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define FRT_HD __host__ __device__ __forceinline__
#else
#define FRT_HD static inline
#endif

namespace frt_dm {

FRT_HD uint64_t bits_of(double x) {
  uint64_t u;
  memcpy(&u, &x, 8);
  return u;
}
FRT_HD double double_of(uint64_t u) {
  double x;
  memcpy(&x, &u, 8);
  return x;
}
FRT_HD double fmad(double a, double b, double c) { return fma(a, b, c); }
FRT_HD double inf() { return double_of(0x7ff0000000000000ull); }
FRT_HD double qnan() { return double_of(0x7ff8000000000000ull); }

// 2^k for k in [-1022, 1023], built from the exponent field (exact).
FRT_HD double pow2i(int k) { return double_of((uint64_t)(k + 1023) << 52); }

// p * 2^k, p in [0.5, 2); one rounding at most (only for subnormal results).
FRT_HD double scale2(double p, int k) {
  if (k > 1023) return (p * pow2i(k - 1)) * 2.0;
  if (k >= -1021) return p * pow2i(k);
  return (p * pow2i(k + 600)) * pow2i(-600);
}

}  // namespace frt_dm

// exp(x): Cody-Waite reduction by ln2, degree-13 Taylor polynomial.
FRT_HD double dm_exp(double x) {
  using namespace frt_dm;
  if (!(x == x)) return x + x;
  if (x > 709.782712893384) return inf();
  if (x < -745.1332191019412) return 0.0;
  const double INV_LN2 = 1.4426950408889634;
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double kd = rint(x * INV_LN2);
  double r = fmad(-kd, LN2_HI, x);
  r = fmad(-kd, LN2_LO, r);
  double p = 1.6059043836821613e-10;  // 1/13!
  p = fmad(p, r, 2.08767569878681e-09);
  p = fmad(p, r, 2.505210838544172e-08);
  p = fmad(p, r, 2.755731922398589e-07);
  p = fmad(p, r, 2.7557319223985893e-06);
  p = fmad(p, r, 2.48015873015873e-05);
  p = fmad(p, r, 0.0001984126984126984);
  p = fmad(p, r, 0.001388888888888889);
  p = fmad(p, r, 0.008333333333333333);
  p = fmad(p, r, 0.041666666666666664);
  p = fmad(p, r, 0.16666666666666666);
  p = fmad(p, r, 0.5);
  p = fmad(p, r, 1.0);
  p = fmad(p, r, 1.0);
  return scale2(p, (int)kd);
}

// log(x): reduce to m in [sqrt(1/2), sqrt(2)), log(m) = f - (f^2/2 - s (f^2/2 + R)),
// s = f/(2+f), R = series of 2 atanh(s) - 2 s in z = s^2.
FRT_HD double dm_log(double x) {
  using namespace frt_dm;
  if (!(x == x)) return x + x;
  if (x < 0.0) return qnan();
  if (x == 0.0) return -inf();
  if (x == inf()) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) {
    x *= 18014398509481984.0;  // 2^54
    e = -54;
  }
  uint64_t u = bits_of(x);
  e += (int)((u >> 52) & 0x7ff) - 1023;
  double m = double_of((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  const double LN2_HI = 6.93147180369123816490e-01;
  const double LN2_LO = 1.90821492927058770002e-10;
  double f = m - 1.0;
  double hfsq = 0.5 * f * f;
  double s = f / (2.0 + f);
  double z = s * s;
  double R = 0.08695652173913043;  // 2/23
  R = fmad(R, z, 0.09523809523809523);  // 2/21
  R = fmad(R, z, 0.10526315789473684);  // 2/19
  R = fmad(R, z, 0.11764705882352941);  // 2/17
  R = fmad(R, z, 0.13333333333333333);  // 2/15
  R = fmad(R, z, 0.15384615384615385);  // 2/13
  R = fmad(R, z, 0.18181818181818182);  // 2/11
  R = fmad(R, z, 0.2222222222222222);   // 2/9
  R = fmad(R, z, 0.2857142857142857);   // 2/7
  R = fmad(R, z, 0.4);                  // 2/5
  R = fmad(R, z, 0.6666666666666666);   // 2/3
  R = R * z;
  double dk = (double)e;
  return dk * LN2_HI - ((hfsq - (s * (hfsq + R) + dk * LN2_LO)) - f);
}

FRT_HD double dm_log10(double x) { return dm_log(x) * 0.43429448190325176; }

// pow(x, y) = exp(y log x), with the IEEE special cases the path can meet.
// Where a correctly rounded pow is a single IEEE operation (y = 1, 2, 0.5, or
// an exact power of two) that operation is used, so those cases match libm.
FRT_HD double dm_pow(double x, double y) {
  using namespace frt_dm;
  if (y == 0.0) return 1.0;
  if (x == 1.0) return 1.0;
  if (!(x == x) || !(y == y)) return x + y;
  if (y == 1.0) return x;
  if (y == 2.0) return x * x;
  if (y == 0.5 && x >= 0.0) return (x == 0.0) ? 0.0 : sqrt(x);
  if (x == 2.0 && rint(y) == y && y >= -1022.0 && y <= 1023.0) return pow2i((int)y);
  if (x == 0.0) return (y > 0.0) ? 0.0 : inf();
  if (x > 0.0) {
    if (x == inf()) return (y > 0.0) ? inf() : 0.0;
    return dm_exp(y * dm_log(x));
  }
  // x < 0: defined for integer y only
  if (rint(y) != y) return qnan();
  double r = dm_exp(y * dm_log(-x));
  double half = 0.5 * y;
  return (rint(half) != half) ? -r : r;
}

// x^n for small integer n >= 0 by repeated multiplication, left to right.
FRT_HD double dm_powi(double x, int n) {
  double r = 1.0;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

namespace frt_dm {
// sin and cos of r, |r| <= pi/4 (+ rounding), Taylor to r^17 / r^16.
FRT_HD void sincos_kernel(double r, double* s, double* c) {
  double z = r * r;
  double ps = 2.8114572543455206e-15;  // 1/17!
  ps = fmad(ps, z, -7.647163731819816e-13);
  ps = fmad(ps, z, 1.6059043836821613e-10);
  ps = fmad(ps, z, -2.505210838544172e-08);
  ps = fmad(ps, z, 2.7557319223985893e-06);
  ps = fmad(ps, z, -0.0001984126984126984);
  ps = fmad(ps, z, 0.008333333333333333);
  ps = fmad(ps, z, -0.16666666666666666);
  *s = fmad(ps * z, r, r);
  double pc = 4.779477332387385e-14;  // 1/16!
  pc = fmad(pc, z, -1.1470745597729725e-11);
  pc = fmad(pc, z, 2.08767569878681e-09);
  pc = fmad(pc, z, -2.755731922398589e-07);
  pc = fmad(pc, z, 2.48015873015873e-05);
  pc = fmad(pc, z, -0.001388888888888889);
  pc = fmad(pc, z, 0.041666666666666664);
  pc = fmad(pc, z, -0.5);
  *c = fmad(pc, z, 1.0);
}
}  // namespace frt_dm

// sin and cos for |x| up to ~1e5 (two-part Cody-Waite reduction by pi/2).
FRT_HD void dm_sincos(double x, double* sn, double* cs) {
  using namespace frt_dm;
  if (!(x == x) || x == inf() || x == -inf()) {
    *sn = qnan();
    *cs = qnan();
    return;
  }
  const double TWO_OVER_PI = 0.6366197723675814;
  const double PIO2_HI = 1.5707963267948966;
  const double PIO2_LO = 6.123233995736766e-17;
  double kd = rint(x * TWO_OVER_PI);
  double r = fmad(-kd, PIO2_HI, x);
  r = fmad(-kd, PIO2_LO, r);
  double s, c;
  sincos_kernel(r, &s, &c);
  int q = ((int)kd) & 3;
  if (q == 0) {
    *sn = s; *cs = c;
  } else if (q == 1) {
    *sn = c; *cs = -s;
  } else if (q == 2) {
    *sn = -s; *cs = -c;
  } else {
    *sn = -c; *cs = s;
  }
}
FRT_HD double dm_sin(double x) { double s, c; dm_sincos(x, &s, &c); return s; }
FRT_HD double dm_cos(double x) { double s, c; dm_sincos(x, &s, &c); return c; }

namespace frt_dm {
// atan(t) for t in [0, 1]: nearest breakpoint c = i/8, atan(t) = atan(c) + atan(u),
// u = (t - c) / (1 + t c), |u| <= 1/16, Taylor to u^15.
FRT_HD double atan01(double t) {
  const double ATAN_I8[9] = {0.0,
                             0.12435499454676144,
                             0.24497866312686414,
                             0.35877067027057225,
                             0.4636476090008061,
                             0.5585993153435624,
                             0.6435011087932844,
                             0.7188299996216245,
                             0.7853981633974483};
  int i = (int)(t * 8.0 + 0.5);
  if (!(i >= 0)) i = 0;
  if (i > 8) i = 8;
  double c = (double)i * 0.125;
  double u = (t - c) / (1.0 + t * c);
  double z = u * u;
  double p = -0.06666666666666667;  // -1/15
  p = fmad(p, z, 0.07692307692307693);   // 1/13
  p = fmad(p, z, -0.09090909090909091);  // -1/11
  p = fmad(p, z, 0.1111111111111111);    // 1/9
  p = fmad(p, z, -0.14285714285714285);  // -1/7
  p = fmad(p, z, 0.2);                   // 1/5
  p = fmad(p, z, -0.3333333333333333);   // -1/3
  return ATAN_I8[i] + fmad(p * z, u, u);
}
}  // namespace frt_dm

// atan2(y, x) with the IEEE sign and quadrant conventions.
FRT_HD double dm_atan2(double y, double x) {
  using namespace frt_dm;
  const double PI = 3.141592653589793;
  const double PIO2 = 1.5707963267948966;
  if (!(x == x) || !(y == y)) return x + y;
  double ax = fabs(x), ay = fabs(y);
  double a;
  if (ay == 0.0) {
    a = (signbit(x)) ? PI : 0.0;
  } else if (ax == 0.0) {
    a = PIO2;
  } else if (ax == inf() && ay == inf()) {
    a = (signbit(x)) ? 2.356194490192345 : 0.7853981633974483;
  } else if (ax == inf()) {
    a = (signbit(x)) ? PI : 0.0;
  } else if (ay == inf()) {
    a = PIO2;
  } else {
    if (ay > ax) {
      a = PIO2 - atan01(ax / ay);
    } else {
      a = atan01(ay / ax);
    }
    if (signbit(x)) a = PI - a;
  }
  return signbit(y) ? -a : a;
}

// Single-precision entry points: evaluate in double, round once to float.
FRT_HD float dm_expf(float x) { return (float)dm_exp((double)x); }
FRT_HD float dm_sinf(float x) { return (float)dm_sin((double)x); }
FRT_HD float dm_cosf(float x) { return (float)dm_cos((double)x); }
