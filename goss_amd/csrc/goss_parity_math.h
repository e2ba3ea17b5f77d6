// FP64 exp, log, pow and sqrt made only of IEEE-754 additions,
// subtractions, multiplications and divisions, floor, and exponent-field
// manipulation, for the parity build (GOSS_PARITY_MATH).
//
// The host library (glibc) and the device library (ROCm ocml) round their
// elementary functions differently in the last bit, and the implicit and
// adaptive solvers turn such differences into different discrete decisions.
// Compiled without FMA contraction (-ffp-contract=off), these routines give
// the same bits on the host and on the GPU, so a host build and a device build
// of the same solver code can be compared bit for bit.  They are accurate to a
// few ulp; the product does not use them.
//
// Methods:
//  exp  -- Cody and Waite, "Software Manual for the Elementary Functions"
//          (1980): x = k ln2 + r with ln2 split into a 32-bit head and a
//          tail, |r| <= ln2/2; exp(r) by its Taylor series to degree 14
//          (truncation below 1e-18 relative); scaling by 2^k through the
//          exponent field.
//  log  -- x = 2^e m with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
//          s = (m - 1)/(m + 1), |s| <= 0.1716, by the odd series
//          2 (s + s^3/3 + ... + s^23/23) (truncation below 1e-19); plus
//          e ln2 with the same split constant (the classical reduction of
//          Cody and Waite, ch. 5).
//  pow  -- exp(y log x), with the special cases of C99 pow for x <= 0.
//  sqrt -- an exponent-halving first guess and six Newton steps
//          y <- (y + x/y)/2 (Heron), converged to the last bits.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif
#ifndef GOSS_HD
#if defined(__HIPCC__)
#define GOSS_HD __host__ __device__
#define GOSS_INLINE __attribute__((always_inline)) inline
#else
#define GOSS_HD
#define GOSS_INLINE inline
#endif
#endif

namespace goss_pm {

GOSS_HD GOSS_INLINE uint64_t bits(const double x)
{
  uint64_t u;
  std::memcpy(&u, &x, sizeof(u));
  return u;
}

GOSS_HD GOSS_INLINE double from_bits(const uint64_t u)
{
  double x;
  std::memcpy(&x, &u, sizeof(x));
  return x;
}

// 2^k for -1022 <= k <= 1023.
GOSS_HD GOSS_INLINE double pow2i(const int k) { return from_bits(uint64_t(k + 1023) << 52); }

// x * 2^k for any int k in [-2200, 2200], with one rounding at most (the last multiply).
GOSS_HD GOSS_INLINE double scale2(double x, int k)
{
  if (k > 1023) {
    x *= pow2i(1023);
    k -= 1023;
    if (k > 1023) {
      x *= pow2i(1023);
      k -= 1023;
      if (k > 1023) k = 1023;
    }
  } else if (k < -1022) {
    x *= pow2i(-1022 + 53);  // keep the intermediate normal
    k += 1022 - 53;
    if (k < -1022) {
      x *= pow2i(-1022 + 53);
      k += 1022 - 53;
      if (k < -1022) k = -1022;
    }
  }
  return x * pow2i(k);
}

constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 0x3FE62E42FEE00000: 32 significant bits
constexpr double kLn2Lo = 1.90821492927058770002e-10;  // ln 2 - kLn2Hi
constexpr double kInvLn2 = 1.44269504088896338700e+00;

GOSS_HD GOSS_INLINE double exp(const double x)
{
  if (x != x) return x;
  if (x > 709.782712893384) return from_bits(0x7ff0000000000000ull);
  if (x < -745.1332191019412) return 0.0;
  const double k = std::floor(x * kInvLn2 + 0.5);
  // k * kLn2Hi is exact for |k| < 2^21 (kLn2Hi has 32 significant bits).
  const double r = (x - k * kLn2Hi) - k * kLn2Lo;
  double p = 1.0 / 87178291200.0;  // 1/14!
  p = p * r + 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return scale2(p, int(k));
}

GOSS_HD GOSS_INLINE double log(double x)
{
  if (x != x || x < 0.0) return from_bits(0x7ff8000000000000ull);
  if (x == 0.0) return -from_bits(0x7ff0000000000000ull);
  if (bits(x) == 0x7ff0000000000000ull) return x;
  int e = 0;
  if (x < 2.2250738585072014e-308) {  // subnormal: make it normal first
    x *= pow2i(54);
    e = -54;
  }
  const uint64_t u = bits(x);
  e += int((u >> 52) & 0x7ff) - 1023;
  double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  double q = 1.0 / 23.0;
  q = q * z + 1.0 / 21.0;
  q = q * z + 1.0 / 19.0;
  q = q * z + 1.0 / 17.0;
  q = q * z + 1.0 / 15.0;
  q = q * z + 1.0 / 13.0;
  q = q * z + 1.0 / 11.0;
  q = q * z + 1.0 / 9.0;
  q = q * z + 1.0 / 7.0;
  q = q * z + 1.0 / 5.0;
  q = q * z + 1.0 / 3.0;
  const double lm = 2.0 * s + 2.0 * s * (z * q);
  const double de = double(e);
  return (de * kLn2Hi + lm) + de * kLn2Lo;
}

GOSS_HD GOSS_INLINE bool is_integer(const double y) { return std::floor(y) == y; }

GOSS_HD GOSS_INLINE double pow(const double x, const double y)
{
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x != x || y != y) return from_bits(0x7ff8000000000000ull);
  if (x == 0.0) return y > 0.0 ? 0.0 : from_bits(0x7ff0000000000000ull);
  if (x < 0.0) {
    if (!is_integer(y)) return from_bits(0x7ff8000000000000ull);
    const double a = goss_pm::exp(y * goss_pm::log(-x));
    const bool odd = std::floor(y * 0.5) * 2.0 != y;
    return odd ? -a : a;
  }
  return goss_pm::exp(y * goss_pm::log(x));
}

GOSS_HD GOSS_INLINE double sqrt(const double x)
{
  if (x != x || x < 0.0) return from_bits(0x7ff8000000000000ull);
  if (x == 0.0 || bits(x) == 0x7ff0000000000000ull) return x;
  double v = x;
  int e = 0;
  if (v < 2.2250738585072014e-308) {
    v *= pow2i(54);
    e = -54;
  }
  const uint64_t u = bits(v);
  e += int((u >> 52) & 0x7ff) - 1023;
  double m = from_bits((u & 0x000fffffffffffffull) | 0x3ff0000000000000ull);  // [1, 2)
  if (e & 1) {
    m *= 2.0;  // [2, 4), e even
    e -= 1;
  }
  // sqrt(m) for m in [1, 4): linear first guess, then Heron's iteration.
  double y = 0.5 + 0.5 * m * 0.5 + 0.25;
  for (int i = 0; i < 6; ++i) y = 0.5 * (y + m / y);
  return scale2(y, e / 2);
}

}  // namespace goss_pm
