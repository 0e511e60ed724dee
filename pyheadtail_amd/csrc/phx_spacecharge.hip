// 2.5-D transverse Gaussian space charge (spacecharge/spacecharge.py:159-417)
// and the device Faddeeva function behind it.
#include "phx_device.h"
#include "phx_faddeeva.h"
#include "phx_host.h"

namespace {

const WofzCoeffs &wofz_host_coeffs() {
  static const WofzCoeffs c = wofz_coefficients();
  return c;
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_math_wofz(const double *__restrict__ x, const double *__restrict__ y, int64_t n,
                double *__restrict__ wr, double *__restrict__ wi, WofzCoeffs c) {
  const int64_t i = (int64_t)blockIdx.x * PHX_BLOCK + threadIdx.x;
  if (i >= n) return;
  double a, b;
  phx_wofz(c, x[i], y[i], a, b);
  wr[i] = a;
  wi[i] = b;
}

// TransverseGaussianSpaceCharge._efieldn_mit for x, y >= 0, sig_x > sig_y
// (:275-300): the operation order of the reference's expression.
__device__ __forceinline__ void efieldn_mit(const WofzCoeffs &c, double x, double y, double sx,
                                            double sy, double denom_pre, double &ex_out,
                                            double &ey_out) {
  const double sig_sqrt = sqrt(2.0 * (sx * sx - sy * sy));
  double w1r, w1i, w2r, w2i;
  phx_wofz(c, x / sig_sqrt, y / sig_sqrt, w1r, w1i);
  const double ex = exp(-x * x / (2.0 * sx * sx) + -y * y / (2.0 * sy * sy));
  phx_wofz(c, x * sy / (sx * sig_sqrt), y * sx / (sy * sig_sqrt), w2r, w2i);
  const double denom = denom_pre * sig_sqrt;  // 2 eps0 sqrt(pi) sig_sqrt
  ex_out = (w1i - ex * w2i) / denom;
  ey_out = (w1r - ex * w2r) / denom;
}

// _efieldn_round (:392-399), x, y >= 0.
__device__ __forceinline__ void efieldn_round(double x, double y, double sr, double round_pre,
                                              double &ex_out, double &ey_out) {
  const double r2 = x * x + y * y;
  const double amplitude = (1.0 - exp(-r2 / (2.0 * sr * sr))) / (round_pre * r2);
  ex_out = x * amplitude;
  ey_out = y * amplitude;
}

__device__ __forceinline__ double sign_of(double v) {
  return v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : (v == 0.0 ? 0.0 : v));
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_spacecharge(double *__restrict__ x, double *__restrict__ xp, double *__restrict__ y,
                  double *__restrict__ yp, const double *__restrict__ z, int64_t n,
                  const int32_t *__restrict__ idx, const phx_slice_state *state,
                  phx_sc_slices sl, phx_spacecharge p, WofzCoeffs c) {
  const double tail = state->tail, width = state->width;
  const int S = sl.n_slices;
  for (int64_t j = (int64_t)blockIdx.x * PHX_BLOCK + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * PHX_BLOCK) {
    int32_t s;
    if (idx) {
      s = idx[j];
    } else {
      const double f = slice_floor(z[j], tail, width);
      s = (f >= 0.0 && f < (double)S) ? (int32_t)f : -1;
    }
    if (s < 0 || s >= S) continue;
    // Nlambda_i = lambda_bins(smoothen=False) / slice_widths
    const double nl = ((double)sl.counts[s] * p.charge_per_mp) / (sl.z_bins[s + 1] - sl.z_bins[s]);
    const double sx = sl.sigma_x[s], sy = sl.sigma_y[s];
    const double dx = x[j] - sl.mean_x[s], dy = y[j] - sl.mean_y[s];
    const double ax = fabs(dx), ay = fabs(dy);
    double ex, ey;
    // add_sigma_check (:401-417): np.allclose(sig_y, sig_x, rtol, atol),
    // np.allclose(sig_y, 0, rtol, atol), sig_x < sig_y
    const bool round = p.sig_check && fabs(sy - sx) <= p.atol + p.rtol * fabs(sx);
    if (round) {
      if (fabs(sy - 0.0) <= p.atol + p.rtol * 0.0) {
        ex = ey = 0.0;
      } else {
        efieldn_round(ax, ay, sx, p.round_pre, ex, ey);
      }
    } else if (p.sig_check && sx < sy) {
      efieldn_mit(c, ay, ax, sy, sx, p.denom_pre, ey, ex);
    } else {
      efieldn_mit(c, ax, ay, sx, sy, p.denom_pre, ex, ey);
    }
    // get_efieldn (:233-246), then the kick (:209-224)
    ex = fabs(ex) * sign_of(dx);
    ey = fabs(ey) * sign_of(dy);
    xp[j] = xp[j] + (ex * nl) * p.prefactor;
    yp[j] = yp[j] + (ey * nl) * p.prefactor;
  }
}

}  // namespace

extern "C" int phx_math_wofz(const double *x, const double *y, int64_t n, double *wr,
                             double *wi, phx_stream_t stream) {
  if (!x || !y || !wr || !wi || n < 0)
    return phx::fail(PHX_E_INVALID, "phx_math_wofz: bad arguments");
  if (n == 0) return PHX_OK;
  const int64_t blocks = (n + PHX_BLOCK - 1) / PHX_BLOCK;
  if (blocks > 0x7fffffffLL) return phx::fail(PHX_E_LIMIT, "phx_math_wofz: n too large");
  k_math_wofz<<<dim3((unsigned)blocks), PHX_BLOCK, 0, phx::as_stream(stream)>>>(
      x, y, n, wr, wi, wofz_host_coeffs());
  PHX_CHECK_LAUNCH("phx_math_wofz");
  return PHX_OK;
}

extern "C" int phx_spacecharge_kick(double *const coords[6], int64_t n, const int32_t *idx,
                                    const phx_slice_state *state, const phx_sc_slices *slices,
                                    const phx_spacecharge *p, phx_stream_t stream) {
  if (!coords || !state || !slices || !p)
    return phx::fail(PHX_E_INVALID, "phx_spacecharge_kick: null pointer");
  for (int i = 0; i < 4; ++i)
    if (!coords[i]) return phx::fail(PHX_E_INVALID, "phx_spacecharge_kick: coordinate %d null", i);
  if (!idx && !coords[4])
    return phx::fail(PHX_E_INVALID, "phx_spacecharge_kick: neither idx nor z");
  const phx_sc_slices &s = *slices;
  if (s.n_slices < 1 || !s.counts || !s.z_bins || !s.mean_x || !s.mean_y || !s.sigma_x ||
      !s.sigma_y)
    return phx::fail(PHX_E_INVALID, "phx_spacecharge_kick: incomplete slice data");
  if (n < 0) return phx::fail(PHX_E_INVALID, "phx_spacecharge_kick: n < 0");
  if (n == 0) return PHX_OK;
  const int64_t blocks = phx::n_chunks(n);
  k_spacecharge<<<dim3((unsigned)(blocks < 4096 ? blocks : 4096)), PHX_BLOCK, 0,
                  phx::as_stream(stream)>>>(coords[0], coords[1], coords[2], coords[3],
                                            coords[4], n, idx, state, s, *p,
                                            wofz_host_coeffs());
  PHX_CHECK_LAUNCH("phx_spacecharge_kick");
  return PHX_OK;
}
