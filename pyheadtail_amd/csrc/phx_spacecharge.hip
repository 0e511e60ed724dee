// 2.5-D transverse Gaussian space charge (spacecharge/spacecharge.py:159-417)
// and the device Faddeeva function behind it.
#include "phx_host.h"
#include "phx_sc.h"

const WofzCoeffs &phx_wofz_host_coeffs() {
  static const WofzCoeffs c = wofz_coefficients();
  return c;
}

namespace {

const WofzCoeffs &wofz_host_coeffs() { return phx_wofz_host_coeffs(); }

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
    const double nl = sc_line_density(sl.counts[s], p.charge_per_mp, sl.z_bins[s],
                                      sl.z_bins[s + 1]);
    double xpj = xp[j], ypj = yp[j];
    sc_kick_particle(c, p, sl.mean_x[s], sl.mean_y[s], sl.sigma_x[s], sl.sigma_y[s], nl, x[j],
                     y[j], xpj, ypj);
    xp[j] = xpj;
    yp[j] = ypj;
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
