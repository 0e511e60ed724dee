// 2.5-D transverse Gaussian space charge (spacecharge/spacecharge.py:
// 159-417): the Bassetti-Erskine field of a slice's Gaussian and the kick of
// one particle, shared by the element kernel (phx_spacecharge.hip) and the
// fused engine's space-charge stages, which therefore kick with the same
// arithmetic.
#pragma once

#include "phx_device.h"
#include "phx_faddeeva.h"

// The Faddeeva coefficients computed once on the host (phx_spacecharge.hip).
const WofzCoeffs &phx_wofz_host_coeffs();

// TransverseGaussianSpaceCharge._efieldn_mit for x, y >= 0, sig_x > sig_y
// (:275-300): the operation order of the reference's expression, for N
// points in lockstep.
template <int N>
__device__ __forceinline__ void sc_efieldn_mit_n(const WofzCoeffs &c, const double (&x)[N],
                                                 const double (&y)[N], const double (&sx)[N],
                                                 const double (&sy)[N], double denom_pre,
                                                 double (&ex_out)[N], double (&ey_out)[N]) {
  double ss[N], u[N], v[N], w1r[N], w1i[N], w2r[N], w2i[N], ex[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    ss[k] = sqrt(2.0 * (sx[k] * sx[k] - sy[k] * sy[k]));
    u[k] = x[k] / ss[k];
    v[k] = y[k] / ss[k];
  }
  phx_wofz_n<N>(c, u, v, w1r, w1i);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    ex[k] = exp(-x[k] * x[k] / (2.0 * sx[k] * sx[k]) + -y[k] * y[k] / (2.0 * sy[k] * sy[k]));
    u[k] = x[k] * sy[k] / (sx[k] * ss[k]);
    v[k] = y[k] * sx[k] / (sy[k] * ss[k]);
  }
  phx_wofz_n<N>(c, u, v, w2r, w2i);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double denom = denom_pre * ss[k];  // 2 eps0 sqrt(pi) sig_sqrt
    ex_out[k] = (w1i[k] - ex[k] * w2i[k]) / denom;
    ey_out[k] = (w1r[k] - ex[k] * w2r[k]) / denom;
  }
}

// _efieldn_round (:392-399), x, y >= 0.
__device__ __forceinline__ void sc_efieldn_round(double x, double y, double sr,
                                                 double round_pre, double &ex_out,
                                                 double &ey_out) {
  const double r2 = x * x + y * y;
  const double amplitude = (1.0 - exp(-r2 / (2.0 * sr * sr))) / (round_pre * r2);
  ex_out = x * amplitude;
  ey_out = y * amplitude;
}

__device__ __forceinline__ double sc_sign_of(double v) {
  return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : (v == 0.0 ? 0.0 : v));
}

// The charge-normalised field E/Q at N particles, each of its own slice:
// offsets (dx, dy) from the slice means, slice standard deviations (sx,
// sy).  add_sigma_check (:401-417): np.allclose(sig_y, sig_x, rtol, atol)
// gives the round-beam field (zero for np.allclose(sig_y, 0)), sig_x <
// sig_y exchanges x and y; get_efieldn (:233-246) evaluates at |dx|, |dy|
// and restores the signs.  The Bassetti-Erskine field is evaluated for
// every particle with its operands exchanged by selects (the particles of a
// wave belong to different slices, so branches would run both forms); the
// round-beam lanes then take their own field.
template <int N>
__device__ __forceinline__ void sc_field_n(const WofzCoeffs &c, const phx_spacecharge &p,
                                           const double (&dx)[N], const double (&dy)[N],
                                           const double (&sx)[N], const double (&sy)[N],
                                           double (&ex)[N], double (&ey)[N]) {
  double a1[N], a2[N], s1[N], s2[N], e1[N], e2[N];
  bool swap[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double ax = fabs(dx[k]), ay = fabs(dy[k]);
    swap[k] = p.sig_check && sx[k] < sy[k];
    a1[k] = swap[k] ? ay : ax;
    a2[k] = swap[k] ? ax : ay;
    s1[k] = swap[k] ? sy[k] : sx[k];
    s2[k] = swap[k] ? sx[k] : sy[k];
  }
  sc_efieldn_mit_n<N>(c, a1, a2, s1, s2, p.denom_pre, e1, e2);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double fx = swap[k] ? e2[k] : e1[k];
    double fy = swap[k] ? e1[k] : e2[k];
    const bool round = p.sig_check && fabs(sy[k] - sx[k]) <= p.atol + p.rtol * fabs(sx[k]);
    if (round) {
      fx = fy = 0.0;
      if (!(fabs(sy[k] - 0.0) <= p.atol + p.rtol * 0.0))
        sc_efieldn_round(fabs(dx[k]), fabs(dy[k]), sx[k], p.round_pre, fx, fy);
    }
    ex[k] = fabs(fx) * sc_sign_of(dx[k]);
    ey[k] = fabs(fy) * sc_sign_of(dy[k]);
  }
}

// The kick of one particle of a slice with means (mx, my), standard
// deviations (sx, sy) and line density nl = Nlambda (lambda_bins / slice
// width): xp += (E_x/Q * Nlambda) * prefactor, yp likewise (:209-224).
__device__ __forceinline__ void sc_kick_particle(const WofzCoeffs &c, const phx_spacecharge &p,
                                                 double mx, double my, double sx, double sy,
                                                 double nl, double x, double y, double &xp,
                                                 double &yp) {
  const double dx[1] = {x - mx}, dy[1] = {y - my}, sxa[1] = {sx}, sya[1] = {sy};
  double ex[1], ey[1];
  sc_field_n<1>(c, p, dx, dy, sxa, sya, ex, ey);
  xp = xp + (ex[0] * nl) * p.prefactor;
  yp = yp + (ey[0] * nl) * p.prefactor;
}

// Nlambda of a slice: lambda_bins(smoothen=False) / slice_widths, i.e.
// (count * charge per macroparticle) / (z_bins[s+1] - z_bins[s]).
__device__ __forceinline__ double sc_line_density(int32_t count, double charge_per_mp,
                                                  double z_left, double z_right) {
  return ((double)count * charge_per_mp) / (z_right - z_left);
}
