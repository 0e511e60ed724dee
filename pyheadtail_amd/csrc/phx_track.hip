// Per-particle element maps as standalone kernels: one pass over the
// coordinates each, PHX_PPT particles per thread with all loads issued
// before the arithmetic.
//   TransverseSegmentMap.track  trackers/transverse_tracking.py:138-196
//   Drift / Kick / RFSystems    trackers/longitudinal_tracking.py:84-203, 556-574
//   LinearMap                   trackers/longitudinal_tracking.py:720-759
#include "phx_device.h"
#include "phx_host.h"

namespace {

inline unsigned grid_for(int64_t n) { return (unsigned)phx::n_chunks(n); }

__global__ void __launch_bounds__(PHX_BLOCK)
    k_transverse(double *__restrict__ x, double *__restrict__ xp, double *__restrict__ y,
                 double *__restrict__ yp, const double *__restrict__ dp, int64_t n, phx_tseg t) {
  const int64_t base = (int64_t)blockIdx.x * PHX_CHUNK + threadIdx.x;
  P6 p[PHX_PPT];
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      p[k].x = x[j];
      p[k].xp = xp[j];
      p[k].y = y[j];
      p[k].yp = yp[j];
      p[k].dp = dp ? dp[j] : 0.0;
      p[k].z = 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      apply_tseg(p[k], t);
      x[j] = p[k].x;
      xp[j] = p[k].xp;
      y[j] = p[k].y;
      yp[j] = p[k].yp;
    }
  }
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_longitudinal(double *__restrict__ x, double *__restrict__ xp, double *__restrict__ y,
                   double *__restrict__ yp, double *__restrict__ z, double *__restrict__ dp,
                   int64_t n, phx_long_seq s, int touches_transverse) {
  const int64_t base = (int64_t)blockIdx.x * PHX_CHUNK + threadIdx.x;
  P6 p[PHX_PPT];
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      p[k].z = z[j];
      p[k].dp = dp[j];
      if (touches_transverse) {
        p[k].x = x[j];
        p[k].xp = xp[j];
        p[k].y = y[j];
        p[k].yp = yp[j];
      } else {
        p[k].x = p[k].xp = p[k].y = p[k].yp = 0.0;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      apply_long_seq(p[k], s);
      z[j] = p[k].z;
      dp[j] = p[k].dp;
      if (touches_transverse) {
        x[j] = p[k].x;
        xp[j] = p[k].xp;
        y[j] = p[k].y;
        yp[j] = p[k].yp;
      }
    }
  }
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_linmap(double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
             double *__restrict__ dp, int64_t n, phx_linmap m) {
  const int64_t base = (int64_t)blockIdx.x * PHX_CHUNK + threadIdx.x;
  P6 p[PHX_PPT];
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      p[k].x = x ? x[j] : 0.0;
      p[k].y = y ? y[j] : 0.0;
      p[k].z = z[j];
      p[k].dp = dp[j];
      p[k].xp = p[k].yp = 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < n) {
      apply_linmap(p[k], m);
      z[j] = p[k].z;
      dp[j] = p[k].dp;
      if (x) x[j] = p[k].x;
      if (y) y[j] = p[k].y;
    }
  }
}

}  // namespace

extern "C" int phx_transverse_segment(double *x, double *xp, double *y, double *yp,
                                      const double *dp, int64_t n, const phx_tseg *p,
                                      phx_stream_t stream) {
  if (!x || !xp || !y || !yp || !p)
    return phx::fail(PHX_E_INVALID, "phx_transverse_segment: null pointer");
  if (n < 0) return phx::fail(PHX_E_INVALID, "phx_transverse_segment: n < 0");
  if (p->n_detuners < 0 || p->n_detuners > PHX_MAX_DETUNERS)
    return phx::fail(PHX_E_LIMIT, "phx_transverse_segment: n_detuners=%d", p->n_detuners);
  for (int i = 0; i < p->n_detuners; ++i) {
    const phx_detuner &d = p->det[i];
    if (d.type != PHX_DET_CHROMA && d.type != PHX_DET_AMPLITUDE)
      return phx::fail(PHX_E_INVALID, "phx_transverse_segment: detuner %d bad type", i);
    if (d.type == PHX_DET_CHROMA && (d.order < 1 || d.order > PHX_MAX_CHROMA))
      return phx::fail(PHX_E_LIMIT, "phx_transverse_segment: chromaticity order %d", d.order);
  }
  const bool needs_dp = p->with_dispersion || p->n_detuners > 0;
  if (needs_dp && !dp) return phx::fail(PHX_E_INVALID, "phx_transverse_segment: dp is null");
  if (n == 0) return PHX_OK;
  k_transverse<<<dim3(grid_for(n)), PHX_BLOCK, 0, phx::as_stream(stream)>>>(x, xp, y, yp, dp, n,
                                                                             *p);
  PHX_CHECK_LAUNCH("phx_transverse_segment");
  return PHX_OK;
}

extern "C" int phx_longitudinal(double *x, double *xp, double *y, double *yp, double *z,
                                double *dp, int64_t n, const phx_long_seq *p,
                                phx_stream_t stream) {
  if (!z || !dp || !p) return phx::fail(PHX_E_INVALID, "phx_longitudinal: null pointer");
  if (n < 0) return phx::fail(PHX_E_INVALID, "phx_longitudinal: n < 0");
  if (p->n_ops < 0 || p->n_ops > PHX_MAX_LONG_OPS)
    return phx::fail(PHX_E_LIMIT, "phx_longitudinal: n_ops=%d", p->n_ops);
  int touches = p->shrink ? 1 : 0;
  for (int i = 0; i < p->n_ops; ++i) {
    if (p->op[i].type != PHX_LONG_DRIFT && p->op[i].type != PHX_LONG_KICK)
      return phx::fail(PHX_E_INVALID, "phx_longitudinal: op %d bad type", i);
    if (p->op[i].type == PHX_LONG_KICK && p->op[i].with_dispersion) touches = 1;
  }
  if (touches && (!x || !xp || !y || !yp))
    return phx::fail(PHX_E_INVALID, "phx_longitudinal: transverse coordinates needed");
  if (n == 0) return PHX_OK;
  k_longitudinal<<<dim3(grid_for(n)), PHX_BLOCK, 0, phx::as_stream(stream)>>>(
      x, xp, y, yp, z, dp, n, *p, touches);
  PHX_CHECK_LAUNCH("phx_longitudinal");
  return PHX_OK;
}

extern "C" int phx_linear_map(double *x, double *y, double *z, double *dp, int64_t n,
                              const phx_linmap *p, phx_stream_t stream) {
  if (!z || !dp || !p || !x || !y)
    return phx::fail(PHX_E_INVALID, "phx_linear_map: null pointer");
  if (n < 0) return phx::fail(PHX_E_INVALID, "phx_linear_map: n < 0");
  if (n == 0) return PHX_OK;
  k_linmap<<<dim3(grid_for(n)), PHX_BLOCK, 0, phx::as_stream(stream)>>>(x, y, z, dp, n, *p);
  PHX_CHECK_LAUNCH("phx_linear_map");
  return PHX_OK;
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_damper(double *__restrict__ xp, double *__restrict__ yp, int64_t n,
             const double *__restrict__ sums, phx_damper p) {
  const double nd = (double)n;
  const double mxp = sums[0] / nd, myp = sums[1] / nd, mx = sums[2] / nd, my = sums[3] / nd;
  const double kxp = p.coef_xp * mxp, kyp = p.coef_yp * myp;
  const double kx = (p.coef_x * mx) / p.beta, ky = (p.coef_y * my) / p.beta;
  const bool hx = p.planes & 1, hy = p.planes & 2, react = p.beta != 0.0;
  for (int64_t i = (int64_t)blockIdx.x * PHX_BLOCK + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * PHX_BLOCK) {
    if (hx) {
      double v = xp[i] - kxp;
      if (react) v = v - kx;
      xp[i] = v;
    }
    if (hy) {
      double v = yp[i] - kyp;
      if (react) v = v - ky;
      yp[i] = v;
    }
  }
}

extern "C" int phx_transverse_damper(double *xp, double *yp, int64_t n, const double *sums,
                                     const phx_damper *p, phx_stream_t stream) {
  if (!p || !sums) return phx::fail(PHX_E_INVALID, "phx_transverse_damper: null pointer");
  if (((p->planes & 1) && !xp) || ((p->planes & 2) && !yp))
    return phx::fail(PHX_E_INVALID, "phx_transverse_damper: null coordinate");
  if (n < 0) return phx::fail(PHX_E_INVALID, "phx_transverse_damper: n < 0");
  if (n == 0 || !(p->planes & 3)) return PHX_OK;
  k_damper<<<dim3(grid_for(n)), PHX_BLOCK, 0, phx::as_stream(stream)>>>(xp, yp, n, sums, *p);
  PHX_CHECK_LAUNCH("phx_transverse_damper");
  return PHX_OK;
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_math_sincos(const double *__restrict__ x, int64_t n, double *__restrict__ s,
                  double *__restrict__ c) {
  const int64_t i = (int64_t)blockIdx.x * PHX_BLOCK + threadIdx.x;
  if (i >= n) return;
  double sv, cv;
  phx_sincos(x[i], &sv, &cv);
  if (s) s[i] = sv;
  if (c) c[i] = cv;
}

extern "C" int phx_math_sincos(const double *x, int64_t n, double *s, double *c,
                               phx_stream_t stream) {
  if (!x || n < 0) return phx::fail(PHX_E_INVALID, "phx_math_sincos: bad arguments");
  if (n == 0) return PHX_OK;
  const int64_t blocks = (n + PHX_BLOCK - 1) / PHX_BLOCK;
  if (blocks > 0x7fffffffLL) return phx::fail(PHX_E_LIMIT, "phx_math_sincos: n too large");
  k_math_sincos<<<dim3((unsigned)blocks), PHX_BLOCK, 0, phx::as_stream(stream)>>>(x, n, s, c);
  PHX_CHECK_LAUNCH("phx_math_sincos");
  return PHX_OK;
}
