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

// The per-slice constants of the field evaluation: the slice's standard
// deviations, the decisions of add_sigma_check (:401-417) and the divisors
// of _efieldn_mit (:275-300) with their correctly rounded reciprocals, so
// that every division by them is div_rcp (the division's bits in three
// operations).
struct ScField {
  double sx, sy;        // slice standard deviations
  double s1, s2;        // the same after the exchange for sig_x < sig_y
  double ss, r_ss;      // sig_sqrt = sqrt(2 (s1^2 - s2^2)) and 1 / ss
  double r_d1, r_d2;    // 1 / (2 s1 s1), 1 / (2 s2 s2)
  double r_m1, r_m2;    // 1 / (s1 ss), 1 / (s2 ss)
  double r_den;         // 1 / (denom_pre ss)
  int32_t swap;         // x and y exchanged (sig_check and sig_x < sig_y)
  int32_t round;        // round-beam field (np.allclose(sig_y, sig_x)); 2: zero
                        // (np.allclose(sig_y, 0) as well)
};

__device__ __forceinline__ ScField sc_field_params(double sx, double sy,
                                                   const phx_spacecharge &p) {
  ScField f;
  f.sx = sx;
  f.sy = sy;
  f.swap = (p.sig_check && sx < sy) ? 1 : 0;
  const bool round = p.sig_check && fabs(sy - sx) <= p.atol + p.rtol * fabs(sx);
  f.round = round ? ((fabs(sy - 0.0) <= p.atol + p.rtol * 0.0) ? 2 : 1) : 0;
  f.s1 = f.swap ? sy : sx;
  f.s2 = f.swap ? sx : sy;
  f.ss = sqrt(2.0 * (f.s1 * f.s1 - f.s2 * f.s2));
  f.r_ss = 1.0 / f.ss;
  f.r_d1 = 1.0 / (2.0 * f.s1 * f.s1);
  f.r_d2 = 1.0 / (2.0 * f.s2 * f.s2);
  f.r_m1 = 1.0 / (f.s1 * f.ss);
  f.r_m2 = 1.0 / (f.s2 * f.ss);
  f.r_den = 1.0 / (p.denom_pre * f.ss);
  return f;
}

// _efieldn_mit (:275-300) at x, y >= 0 with sig_x = s1 > sig_y = s2, for N
// points in lockstep: the operation order of the reference's expression,
// its divisions by the slice constants as div_rcp.  The two Faddeeva
// arguments of every point are formed first and the 2N evaluations run in
// lockstep (independent chains; each evaluation's arithmetic is unchanged).
template <int N>
__device__ __forceinline__ void sc_efieldn_mit_n(const WofzCoeffs &c, const double (&x)[N],
                                                 const double (&y)[N], const ScField (&f)[N],
                                                 double denom_pre, double (&ex_out)[N],
                                                 double (&ey_out)[N]) {
  // [0, N): x / sig_sqrt + i y / sig_sqrt;  [N, 2N): the scaled arguments
  double u[2 * N], v[2 * N], wr[2 * N], wi[2 * N], ex[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double s1 = f[k].s1, s2 = f[k].s2, ss = f[k].ss;
    u[k] = div_rcp(x[k], ss, f[k].r_ss);
    v[k] = div_rcp(y[k], ss, f[k].r_ss);
    ex[k] = exp(div_rcp(-x[k] * x[k], 2.0 * s1 * s1, f[k].r_d1) +
                div_rcp(-y[k] * y[k], 2.0 * s2 * s2, f[k].r_d2));
    u[N + k] = div_rcp(x[k] * s2, s1 * ss, f[k].r_m1);
    v[N + k] = div_rcp(y[k] * s1, s2 * ss, f[k].r_m2);
  }
  phx_wofz_n<2 * N>(c, u, v, wr, wi);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double denom = denom_pre * f[k].ss;  // 2 eps0 sqrt(pi) sig_sqrt
    ex_out[k] = div_rcp(wi[k] - ex[k] * wi[N + k], denom, f[k].r_den);
    ey_out[k] = div_rcp(wr[k] - ex[k] * wr[N + k], denom, f[k].r_den);
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

// The charge-normalised field E/Q at N particles, each of its own slice
// (offsets dx, dy from the slice means; the slice constants f): the
// Bassetti-Erskine field evaluated for every particle with its operands
// exchanged by selects (the particles of a wave belong to different
// slices, so branches would run both forms), the round-beam lanes then
// taking their own field; get_efieldn (:233-246) evaluates at |dx|, |dy|
// and restores the signs.
template <int N>
__device__ __forceinline__ void sc_field_n(const WofzCoeffs &c, const phx_spacecharge &p,
                                           const double (&dx)[N], const double (&dy)[N],
                                           const ScField (&f)[N], double (&ex)[N],
                                           double (&ey)[N]) {
  double a1[N], a2[N], e1[N], e2[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double ax = fabs(dx[k]), ay = fabs(dy[k]);
    a1[k] = f[k].swap ? ay : ax;
    a2[k] = f[k].swap ? ax : ay;
  }
  sc_efieldn_mit_n<N>(c, a1, a2, f, p.denom_pre, e1, e2);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    double fx = f[k].swap ? e2[k] : e1[k];
    double fy = f[k].swap ? e1[k] : e2[k];
    if (f[k].round) {
      fx = fy = 0.0;
      if (f[k].round == 1)
        sc_efieldn_round(fabs(dx[k]), fabs(dy[k]), f[k].sx, p.round_pre, fx, fy);
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
  const double dx[1] = {x - mx}, dy[1] = {y - my};
  const ScField f[1] = {sc_field_params(sx, sy, p)};
  double ex[1], ey[1];
  sc_field_n<1>(c, p, dx, dy, f, ex, ey);
  xp = xp + (ex[0] * nl) * p.prefactor;
  yp = yp + (ey[0] * nl) * p.prefactor;
}

// Nlambda of a slice: lambda_bins(smoothen=False) / slice_widths, i.e.
// (count * charge per macroparticle) / (z_bins[s+1] - z_bins[s]).
__device__ __forceinline__ double sc_line_density(int32_t count, double charge_per_mp,
                                                  double z_left, double z_right) {
  return ((double)count * charge_per_mp) / (z_right - z_left);
}
