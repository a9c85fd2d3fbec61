// gw_math.h — FP64 logarithms for the van Genuchten closures.
//
// The closures are bound by vector instructions on the GPU, and the device
// library's FP64 log and log1p (extended-precision, 98 and 135 vector
// instructions on gfx950) are most of them.  These take ~30 and ~45 for an
// error of a few ulp (checked against long-double libm on the host by
// tests/test_device_math.py), which the closures' 1e-10 parity leaves far
// behind.  Header-only and host-compilable, so the same code is tested on
// the CPU.
//
// log(x): x = m 2^e with m in [1/sqrt2, sqrt2); with f = m - 1 and
// s = f / (2 + f), log m = 2 atanh s = 2 s (1 + s^2/3 + s^4/5 + ... ), the
// argument reduction and series of Cody and Waite (Software Manual for the
// Elementary Functions, 1980), here with the plain Taylor coefficients
// 1/(2k+1): |s| <= 0.1716, so ten terms leave a tail below 2^-58 of the
// result.  With 2 s = f - s f the sum is f - (f^2/2 - s (f^2/2 + R)),
// R = 2 s^2 (1/3 + s^2/5 + ...), so the rounding of s only touches terms
// of order f^2.  log x = e ln2_hi + (e ln2_lo + log m), ln2 split so that
// e ln2_hi is exact for every exponent.  log(0) = -inf, log(+inf) = +inf,
// log(x < 0) and log(NaN) = NaN.
//
// log1p(x) = log(u) x / (u - 1), u = 1 + x (x when u == 1): Goldberg's
// correction for the rounding of u (What Every Computer Scientist Should
// Know About Floating-Point Arithmetic, 1991, theorem 4).
#pragma once

#include <math.h>

#ifndef __HIPCC__
#define GW_MATH_FN static inline
#else
#define GW_MATH_FN __host__ __device__ __forceinline__
#endif

namespace frehg {

GW_MATH_FN double gw_log(double x) {
  const double kLn2Hi = 6.93147180369123816490e-01;  // 0x3fe62e42fee00000: 32 bits
  const double kLn2Lo = 1.9082149292705877e-10;      // ln2 - kLn2Hi
  const double kSqrtHalf = 0.70710678118654752440;
  int e = 0;
  double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  if (m < kSqrtHalf) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;  // exact
  const double s = f / (2.0 + f);
  const double z = s * s;
  double p = 1.0 / 19.0;
  p = fma(p, z, 1.0 / 17.0);
  p = fma(p, z, 1.0 / 15.0);
  p = fma(p, z, 1.0 / 13.0);
  p = fma(p, z, 1.0 / 11.0);
  p = fma(p, z, 1.0 / 9.0);
  p = fma(p, z, 1.0 / 7.0);
  p = fma(p, z, 1.0 / 5.0);
  p = fma(p, z, 1.0 / 3.0);
  // 2 s (1 + z p) written as f - (f^2/2 - s (f^2/2 + 2 z p)) (2 s = f - s f):
  // the rounding of s then only reaches the small correction terms
  const double hfsq = 0.5 * f * f;
  const double lm = f - (hfsq - s * (hfsq + 2.0 * z * p));
  const double ed = static_cast<double>(e);
  double r = fma(ed, kLn2Hi, fma(ed, kLn2Lo, lm));
  if (!(x > 0.0)) r = (x == 0.0) ? -INFINITY : NAN;
  if (x == INFINITY) r = INFINITY;
  return r;
}

GW_MATH_FN double gw_log1p(double x) {
  const double u = 1.0 + x;
  if (u == 1.0) return x;
  if (u == INFINITY) return INFINITY;
  return gw_log(u) * (x / (u - 1.0));
}

}  // namespace frehg
