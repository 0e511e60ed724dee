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
#pragma once

#include <hip/hip_runtime.h>

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

// w(x + i y), y >= 0: real and imaginary parts.
__device__ __forceinline__ void phx_wofz(const WofzCoeffs &c, double x, double y, double &wr,
                                         double &wi) {
  const double L = c.L;
  // L - i z = (L + y) - i x,  L + i z = (L - y) + i x
  const double dr = L + y, di = -x;
  const double den = dr * dr + di * di;
  // q = 1 / (L - i z) = (dr - i di) / den
  const double qr = dr / den, qi = -di / den;
  // Z = (L + i z) q
  const double nr = L - y, ni = x;
  const double Zr = nr * qr - ni * qi, Zi = nr * qi + ni * qr;
  double pr = c.a[0], pi = 0.0;
#pragma unroll 8
  for (int n = 1; n < PHX_WOFZ_N; ++n) {
    const double tr = fma(pr, Zr, fma(-pi, Zi, c.a[n]));
    pi = fma(pr, Zi, pi * Zr);
    pr = tr;
  }
  // w = 2 p q^2 + q / sqrt(pi)
  const double q2r = qr * qr - qi * qi, q2i = 2.0 * qr * qi;
  const double isp = 0.56418958354775628695;  // 1/sqrt(pi)
  wr = 2.0 * (pr * q2r - pi * q2i) + isp * qr;
  wi = 2.0 * (pr * q2i + pi * q2r) + isp * qi;
}
