// FP64 sine and cosine for the particle maps.
//
// The phases of the maps are moderate: the RF phase h*2*pi*z/C stays within
// a few radians of the bucket, the betatron phase 2*pi*(Q + dQ) within a few
// hundred radians.  For |x| < 2^20 the argument is reduced by Cody and
// Waite's method with pi/2 split into P1 + P2 + P3: P1 and P2 carry 33
// significant bits, so k*P1 and k*P2 are exact for |k| < 2^20 and x - k*P1
// is exact (both are multiples of ulp(x) there); the remainder
// r = x - k*pi/2, |r| <~ pi/4, is then within ~1 ulp.  sin(r) and cos(r)
// are their Taylor series to r^17 and r^16 (truncation below 1e-19 at
// pi/4), evaluated by Horner's rule in r^2, and the quadrant k mod 4 picks
// and signs them.  Larger arguments go to the device library through calls
// kept out of line that return by value (an output pointer would put the
// results, and the callers' every launch, on scratch memory).
//
// The constants: P1, P2 are pi/2 truncated to 33 and the next 33 bits, P3
// the rest rounded to double (pi from Machin's formula to 80 digits);
// 2/pi rounded to double; the coefficients are (-1)^k/(2k+1)! and
// (-1)^k/(2k)! rounded to double.
#pragma once

#include <hip/hip_runtime.h>

namespace phx_math {

constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_1 = 0x1.921fb54400000p+0;
constexpr double PIO2_2 = 0x1.0b4611a600000p-34;
constexpr double PIO2_3 = 0x1.3198a2e037073p-69;
constexpr double REDUCE_MAX = 0x1p20;

// sin: r + r^3 * (S1 + r^2 (S2 + ...)),  S_k = (-1)^k / (2k+1)!
constexpr double S1 = -1.0 / 6.0;
constexpr double S2 = 1.0 / 120.0;
constexpr double S3 = -1.0 / 5040.0;
constexpr double S4 = 1.0 / 362880.0;
constexpr double S5 = -1.0 / 39916800.0;
constexpr double S6 = 1.0 / 6227020800.0;
constexpr double S7 = -1.0 / 1307674368000.0;
constexpr double S8 = 1.0 / 355687428096000.0;
// cos: 1 + r^2 * (C1 + r^2 (C2 + ...)),  C_k = (-1)^k / (2k)!
constexpr double C1 = -0.5;
constexpr double C2 = 1.0 / 24.0;
constexpr double C3 = -1.0 / 720.0;
constexpr double C4 = 1.0 / 40320.0;
constexpr double C5 = -1.0 / 3628800.0;
constexpr double C6 = 1.0 / 479001600.0;
constexpr double C7 = -1.0 / 87178291200.0;
constexpr double C8 = 1.0 / 20922789888000.0;

// sin and cos of r + r_lo, |r| <= ~pi/4, with r_lo the tail of the reduced
// argument: sin(r + r_lo) ~ sin(r) + r_lo cos(r), cos(r + r_lo) ~
// cos(r) - r_lo sin(r), both folded in before the last rounding.
__host__ __device__ __forceinline__ void kernel_sincos(double r, double r_lo, double &s,
                                                       double &c) {
  const double z = r * r;
  double ps = S8;
  ps = fma(ps, z, S7);
  ps = fma(ps, z, S6);
  ps = fma(ps, z, S5);
  ps = fma(ps, z, S4);
  ps = fma(ps, z, S3);
  ps = fma(ps, z, S2);
  ps = fma(ps, z, S1);
  double pc = C8;
  pc = fma(pc, z, C7);
  pc = fma(pc, z, C6);
  pc = fma(pc, z, C5);
  pc = fma(pc, z, C4);
  pc = fma(pc, z, C3);
  pc = fma(pc, z, C2);
  pc = fma(pc, z, C1);
  // 1 + z*pc - r*r_lo: the tail enters through the low part of z*pc
  const double zpc = z * pc;
  const double zpc_lo = fma(z, pc, -zpc) - r * r_lo;
  c = 1.0 + (zpc + zpc_lo);
  s = r + fma(r * z, ps, r_lo * (1.0 + zpc));
}

// Argument reduction: x - k*pi/2 = r + r_lo and the quadrant k mod 4.
// x - k*P1 is exact; the subtraction of k*P2 (exact product) is done as an
// exact two-sum, k*P3 is a correction far below ulp(r).
__host__ __device__ __forceinline__ double reduce(double x, double &r_lo, int &quadrant) {
  const double k = rint(x * TWO_OVER_PI);
  const double a = fma(-k, PIO2_1, x);
  const double b = k * PIO2_2;
  const double w = k * PIO2_3;
  const double hi = a - b;
  const double bv = a - hi;  // two-sum of a and -b
  const double lo = (a - (hi + bv)) + (bv - b);
  const double r = hi - w;
  r_lo = ((hi - r) - w) + lo;
  quadrant = ((int)k) & 3;
  return r;
}

__device__ __noinline__ double cos_library(double x) { return ::cos(x); }

__device__ __noinline__ double sin_library(double x) { return ::sin(x); }

}  // namespace phx_math

// sin(x) and cos(x); within ~1 ulp for |x| < 2^20, the device library's
// values beyond (and for NaN / infinity).
__device__ __forceinline__ void phx_sincos(double x, double *sp, double *cp) {
  if (!(fabs(x) < phx_math::REDUCE_MAX)) {
    *sp = phx_math::sin_library(x);
    *cp = phx_math::cos_library(x);
    return;
  }
  int q;
  double r_lo;
  const double r = phx_math::reduce(x, r_lo, q);
  double s, c;
  phx_math::kernel_sincos(r, r_lo, s, c);
  // quadrant q: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s)
  const double sa = (q & 1) ? c : s;
  const double ca = (q & 1) ? s : c;
  *sp = (q & 2) ? -sa : sa;
  *cp = ((q + 1) & 2) ? -ca : ca;
}

__device__ __forceinline__ double phx_sin(double x) {
  if (!(fabs(x) < phx_math::REDUCE_MAX)) return phx_math::sin_library(x);
  int q;
  double r_lo;
  const double r = phx_math::reduce(x, r_lo, q);
  double s, c;
  phx_math::kernel_sincos(r, r_lo, s, c);
  const double sa = (q & 1) ? c : s;
  return (q & 2) ? -sa : sa;
}
