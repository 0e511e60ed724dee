// The Faddeeva function w(z) = exp(-z^2) erfc(-i z) for Im z >= 0, FP64, by
// J. A. C. Weideman's rational approximation (SIAM J. Numer. Anal. 31
// (1994) 1497-1518):
//   w(z) ~ 2 p(Z) / (L - i z)^2 + (1/sqrt(pi)) / (L - i z),
//   Z = (L + i z) / (L - i z),  L = sqrt(N / sqrt(2)),
// with p the degree N-1 polynomial whose coefficients are the discrete
// Fourier coefficients of f(t) = exp(-t^2) (L^2 + t^2) sampled at
// t = L tan(theta/2), theta = k pi / M, |k| < M = 2N.  With N = 40 it agrees
// with scipy.special.wofz to about 2e-14 relative in the first quadrant
// (the space-charge field evaluates w at |x| + i |y| only).
//
// p has real coefficients, so p(Z) is evaluated in real arithmetic as the
// remainder of p on division by the real quadratic Z^2 - r Z + s
// (r = 2 Re Z, s = |Z|^2), whose root Z is.  With u_j the coefficient of
// Z^j (a[N-1-j] below) and d = N - 1: c_d = u_d, c_(d-1) = u_(d-1) + r c_d,
// c_j = u_j + r c_(j+1) - s c_(j+2) down to j = 1, and p(Z) = c_1 Z +
// (u_0 - s c_2) (D. E. Knuth, The Art of Computer Programming, vol. 2,
// 4.6.4).  Two FMAs per coefficient instead of a complex multiply-add's
// four operations; its error against scipy.special.wofz is that of the
// complex Horner form (the approximation's own, 2.1e-14 at most on the
// accuracy test's sample).
#pragma once

#include <hip/hip_runtime.h>

#include "phx_device.h"

#include <cmath>

#define PHX_WOFZ_N 40

struct WofzCoeffs {
  double L;
  double a[PHX_WOFZ_N];  // p(Z) = sum_n a[n] Z^(N-1-n)  (highest power first)
};

// Host: the coefficients by a direct DFT of the 2M samples (f padded with a
// leading 0, then rotated by M as np.fft.fftshift does).
inline WofzCoeffs wofz_coefficients() {
  const int N = PHX_WOFZ_N, M = 2 * N, M2 = 2 * M;
  WofzCoeffs c;
  c.L = std::sqrt(N / std::sqrt(2.0));
  double f[2 * 2 * PHX_WOFZ_N];
  f[0] = 0.0;
  for (int k = -M + 1; k < M; ++k) {
    const double t = c.L * std::tan(k * M_PI / M / 2.0);
    f[k + M] = std::exp(-t * t) * (c.L * c.L + t * t);
  }
  double g[2 * 2 * PHX_WOFZ_N];
  for (int i = 0; i < M2; ++i) g[i] = f[(i + M) % M2];
  // real part of the DFT coefficients 1..N, divided by 2M, highest first
  for (int n = 1; n <= N; ++n) {
    double s = 0.0;
    for (int i = 0; i < M2; ++i) s += g[i] * std::cos(2.0 * M_PI * (double)((long)i * n % M2) / M2);
    c.a[N - n] = s / M2;
  }
  return c;
}

// w(x_k + i y_k), y_k >= 0, for N points: real and imaginary parts.  The N
// evaluations run in lockstep (N independent recurrences for the compiler
// to interleave); each point's arithmetic is that of N = 1.
template <int N>
__device__ __forceinline__ void phx_wofz_n(const WofzCoeffs &c, const double (&x)[N],
                                           const double (&y)[N], double (&wr)[N],
                                           double (&wi)[N]) {
  const double L = c.L;
  double qr[N], qi[N], Zr[N], Zi[N], r[N], s[N], c1[N], c2[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // L - i z = (L + y) - i x,  L + i z = (L - y) + i x
    const double dr = L + y[k], di = -x[k];
    const double den = dr * dr + di * di;
    // q = 1 / (L - i z) = (dr - i di) / den: both quotients from one
    // correctly rounded reciprocal (the division's bits, div_rcp)
    const double rden = 1.0 / den;
    qr[k] = div_rcp(dr, den, rden);
    qi[k] = div_rcp(-di, den, rden);
    // Z = (L + i z) q
    const double nr = L - y[k], ni = x[k];
    Zr[k] = nr * qr[k] - ni * qi[k];
    Zi[k] = nr * qi[k] + ni * qr[k];
    // the quadratic with root Z; the two highest coefficients
    r[k] = 2.0 * Zr[k];
    s[k] = fma(Zr[k], Zr[k], Zi[k] * Zi[k]);
    c2[k] = c.a[0];
    c1[k] = fma(r[k], c2[k], c.a[1]);
  }
#pragma unroll
  for (int n = 2; n < PHX_WOFZ_N - 1; ++n) {
    const double an = c.a[n];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double c0 = fma(r[k], c1[k], fma(-s[k], c2[k], an));
      c2[k] = c1[k];
      c1[k] = c0;
    }
  }
  // w = 2 p q^2 + q / sqrt(pi)
  const double isp = 0.56418958354775628695;  // 1/sqrt(pi)
#pragma unroll
  for (int k = 0; k < N; ++k) {
    // p(Z) = c_1 Z + (a_0 - s c_2)
    const double pr = fma(c1[k], Zr[k], fma(-s[k], c2[k], c.a[PHX_WOFZ_N - 1]));
    const double pi = c1[k] * Zi[k];
    const double q2r = qr[k] * qr[k] - qi[k] * qi[k], q2i = 2.0 * qr[k] * qi[k];
    wr[k] = 2.0 * (pr * q2r - pi * q2i) + isp * qr[k];
    wi[k] = 2.0 * (pr * q2i + pi * q2r) + isp * qi[k];
  }
}

// w(x + i y), y >= 0: real and imaginary parts.
__device__ __forceinline__ void phx_wofz(const WofzCoeffs &c, double x, double y, double &wr,
                                         double &wi) {
  const double xs[1] = {x}, ys[1] = {y};
  double a[1], b[1];
  phx_wofz_n<1>(c, xs, ys, a, b);
  wr = a[0];
  wi = b[0];
}
