// Device-side building blocks shared by the standalone element kernels and
// the fused turn engine.  Everything here is FP64 and is compiled with
// -ffp-contract=off so that each arithmetic operation is one IEEE-754
// rounding, in the order the reference's NumPy expressions evaluate them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/phx.h"

#define PHX_BLOCK 256
#define PHX_WAVE 64
#define PHX_PPT 4
#define PHX_CHUNK (PHX_BLOCK * PHX_PPT)
// NumPy's 2.*np.pi as a double
#define PHX_TWO_PI 6.283185307179586
#define PHX_PI 3.141592653589793

struct P6 {
  double x, xp, y, yp, z, dp;
};

__device__ __forceinline__ double p6_get(const P6 &p, int c) {
  switch (c) {
    case 0: return p.x;
    case 1: return p.xp;
    case 2: return p.y;
    case 3: return p.yp;
    case 4: return p.z;
    default: return p.dp;
  }
}

// ---------------------------------------------------------------------------
// Slice index: int32(floor((z - tail) / width)); particles/slicing.py:609-611.
// Returns the floor as a double; "inside" means 0 <= f < n_slices, which is
// false for NaN exactly as the int32 conversion of NaN (INT32_MIN on x86) is
// outside [0, n_slices).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double slice_floor(double z, double tail, double width) {
  return floor((z - tail) / width);
}

__device__ __forceinline__ int32_t x86_cvt_i32(double f) {
  // cvttsd2si semantics: NaN and out-of-range values give INT32_MIN
  if (!(f >= -2147483648.0 && f < 2147483648.0)) return INT32_MIN;
  return (int32_t)f;
}

// ---------------------------------------------------------------------------
// Transverse segment map (trackers/transverse_tracking.py:138-196) with the
// detuners (trackers/detuners.py:64-98 chromaticity, :128-151 amplitude).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void detune(const phx_detuner &d, const P6 &p, double &qx,
                                       double &qy) {
  if (d.type == PHX_DET_CHROMA) {
    // Horner form of (c0 + (c1 + c2*dp)*dp)*dp, the same association as the
    // explicit order-1..3 polynomials and np.polynomial.polyval.
    double ax = d.cx[d.order - 1], ay = d.cy[d.order - 1];
    for (int i = d.order - 2; i >= 0; --i) {
      ax = d.cx[i] + ax * p.dp;
      ay = d.cy[i] + ay * p.dp;
    }
    qx = ax * p.dp;
    qy = ay * p.dp;
  } else {
    // Jx = 1/2*((1+a^2)/b*x**2 + 2*a*x*xp + b*xp**2)
    double jx = 0.5 * ((d.gx * (p.x * p.x) + (d.ax2 * p.x) * p.xp) + d.bx * (p.xp * p.xp));
    double jy = 0.5 * ((d.gy * (p.y * p.y) + (d.ay2 * p.y) * p.yp) + d.by * (p.yp * p.yp));
    qx = (d.dapp_x * jx + d.dapp_xy * jy) / d.p0;
    qy = (d.dapp_y * jy + d.dapp_yx * jx) / d.p0;
  }
}

__device__ __forceinline__ void apply_tseg(P6 &p, const phx_tseg &t) {
  double M00, M01, M10, M11, M22, M23, M32, M33;
  if (t.n_detuners == 0) {
    M00 = t.M00; M01 = t.M01; M10 = t.M10; M11 = t.M11;
    M22 = t.M22; M23 = t.M23; M32 = t.M32; M33 = t.M33;
  } else {
    double dphi_x = t.dQ_x, dphi_y = t.dQ_y;
    for (int i = 0; i < t.n_detuners; ++i) {
      double qx, qy;
      detune(t.det[i], p, qx, qy);
      dphi_x = dphi_x + qx;
      dphi_y = dphi_y + qy;
    }
    dphi_x = dphi_x * PHX_TWO_PI;
    dphi_y = dphi_y * PHX_TWO_PI;
    double sx, cx, sy, cy;
    sincos(dphi_x, &sx, &cx);
    sincos(dphi_y, &sy, &cy);
    M00 = t.I00 * cx + t.J00 * sx;
    M01 = t.I01 * cx + t.J01 * sx;
    M10 = t.I10 * cx + t.J10 * sx;
    M11 = t.I11 * cx + t.J11 * sx;
    M22 = t.I22 * cy + t.J22 * sy;
    M23 = t.I23 * cy + t.J23 * sy;
    M32 = t.I32 * cy + t.J32 * sy;
    M33 = t.I33 * cy + t.J33 * sy;
  }
  if (t.with_dispersion) {
    p.x = p.x - t.D_x_s0 * p.dp;
    p.y = p.y - t.D_y_s0 * p.dp;
  }
  const double x = p.x, xp = p.xp, y = p.y, yp = p.yp;
  p.x = M00 * x + M01 * xp;
  p.xp = M10 * x + M11 * xp;
  p.y = M22 * y + M23 * yp;
  p.yp = M32 * y + M33 * yp;
  if (t.with_dispersion) {
    p.x = p.x + t.D_x_s1 * p.dp;
    p.y = p.y + t.D_y_s1 * p.dp;
  }
}

// ---------------------------------------------------------------------------
// Longitudinal maps (trackers/longitudinal_tracking.py).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void apply_long_seq(P6 &p, const phx_long_seq &s) {
  for (int i = 0; i < s.n_ops; ++i) {
    const phx_long_op &o = s.op[i];
    if (o.type == PHX_LONG_DRIFT) {
      // Drift.track: z = beta_ratio*z - eta(dp)*dp*length   (:107-111)
      p.z = o.beta_ratio * p.z - (o.eta0 * p.dp) * o.length;
    } else {
      // Kick.track_with_dispersion / track_without_dispersion (:185-203)
      if (o.with_dispersion) {
        p.x = p.x - o.D_x * p.dp;
        p.y = p.y - o.D_y * p.dp;
      }
      const double phi = (o.harmonic * ((o.two_pi * p.z) / o.circumference) + o.phi_offset) +
                         o.phi_lock;
      double delta_p = p.dp * o.p0_before;
      delta_p = delta_p + (o.amplitude * sin(phi) - o.p_increment);
      p.dp = delta_p / o.p0_after;
      if (o.with_dispersion) {
        p.x = p.x + o.D_x * p.dp;
        p.y = p.y + o.D_y * p.dp;
      }
    }
  }
  if (s.shrink) {
    p.x = p.x * s.shrink_factor;
    p.xp = p.xp * s.shrink_factor;
    p.y = p.y * s.shrink_factor;
    p.yp = p.yp * s.shrink_factor;
  }
}

__device__ __forceinline__ void apply_linmap(P6 &p, const phx_linmap &m) {
  // LinearMap.track_with_dispersion -> track_without_dispersion (:731-759)
  p.x = p.x - m.D_x * p.dp;
  p.y = p.y - m.D_y * p.dp;
  const double z0 = p.z, dp0 = p.dp;
  p.z = z0 * m.cosQ - (m.longfac * dp0) * m.sinQ;
  p.dp = dp0 * m.cosQ + (z0 / m.longfac) * m.sinQ;
  p.x = p.x + m.D_x * p.dp;
  p.y = p.y + m.D_y * p.dp;
}

__device__ __forceinline__ void apply_program(P6 &p, const phx_turn_program &g) {
  for (int e = 0; e < g.n_el; ++e) {
    const int k = g.el_index[e];
    switch (g.el_type[e]) {
      case PHX_EL_TSEG: apply_tseg(p, g.tseg[k]); break;
      case PHX_EL_LONG: apply_long_seq(p, g.lseq[k]); break;
      default: apply_linmap(p, g.lin[k]); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Wake kick application (impedances/wake_kicks.py:160-361), kicks in order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void apply_wake_kicks(P6 &p, int32_t idx, int n_slices,
                                                 const phx_wake_kick *kd, int n_kicks,
                                                 const double *__restrict__ kicks) {
  if (idx < 0 || idx >= n_slices) return;
  // gather all kick values first so their loads are in flight together
  double kv[PHX_MAX_KICKS];
#pragma unroll
  for (int k = 0; k < PHX_MAX_KICKS; ++k)
    if (k < n_kicks) kv[k] = kicks[(size_t)k * n_slices + idx];
#pragma unroll
  for (int k = 0; k < PHX_MAX_KICKS; ++k) {
    if (k >= n_kicks) break;
    double v = kv[k];
    if (kd[k].multiply == 0) v = v * p.x;
    else if (kd[k].multiply == 2) v = v * p.y;
    if (kd[k].plane == PHX_PLANE_XP) p.xp = p.xp + v;
    else if (kd[k].plane == PHX_PLANE_YP) p.yp = p.yp + v;
    else p.dp = p.dp + v;
  }
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  Within a wave: a fixed shuffle-down tree whose
// result is taken from lane 0; within a block: the wave results combined in
// wave order.  Across chunks: each thread folds a fixed stride of chunks in
// increasing order, then the same block reduction.  For a given particle
// count the association order is fixed, so results are bit-reproducible.
// The operation of slot i is min if bit i of MINM is set, max if bit i of
// MAXM is set, else sum (masks are compile-time so the slots stay in VGPRs).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double red_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ double red_max(double a, double b) { return fmax(a, b); }

template <unsigned long long MINM, unsigned long long MAXM>
__device__ __forceinline__ double red_slot(int i, double a, double b) {
  if ((MINM >> i) & 1ULL) return fmin(a, b);
  if ((MAXM >> i) & 1ULL) return fmax(a, b);
  return a + b;
}

template <unsigned long long MINM, unsigned long long MAXM>
__device__ __forceinline__ double red_ident(int i) {
  if ((MINM >> i) & 1ULL) return INFINITY;
  if ((MAXM >> i) & 1ULL) return -INFINITY;
  return 0.0;
}

// Block-reduce NV slots held by every thread; thread t < NV then holds the
// block result of slot t in the return value (other threads: undefined).
// lds: __shared__ double[NV * NT / PHX_WAVE] for a block of NT threads.
template <int NV, unsigned long long MINM, unsigned long long MAXM, int NT = PHX_BLOCK>
__device__ __forceinline__ double block_reduce_slots(double (&v)[NV], double *lds) {
  constexpr int NW = NT / PHX_WAVE;
  const int lane = threadIdx.x % PHX_WAVE, wave = threadIdx.x / PHX_WAVE;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double r = v[i];
#pragma unroll
    for (int off = PHX_WAVE / 2; off > 0; off >>= 1)
      r = red_slot<MINM, MAXM>(i, r, __shfl_down(r, off, PHX_WAVE));
    if (lane == 0) lds[i * NW + wave] = r;
  }
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x < NV) {
    const int i = threadIdx.x;
    out = lds[i * NW];
    for (int w = 1; w < NW; ++w) {
      const double b = lds[i * NW + w];
      if ((MINM >> i) & 1ULL) out = fmin(out, b);
      else if ((MAXM >> i) & 1ULL) out = fmax(out, b);
      else out = out + b;
    }
  }
  __syncthreads();
  return out;
}

// Reduce per-chunk partials part[c*NV + i] over c < n_chunks into res[i]
// (__shared__, valid on all threads after return).  Called by one block.
template <int NV, unsigned long long MINM, unsigned long long MAXM>
__device__ __forceinline__ void reduce_partials(const double *part, int64_t n_chunks,
                                                double *res, double *lds) {
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) acc[i] = red_ident<MINM, MAXM>(i);
  for (int64_t c = threadIdx.x; c < n_chunks; c += PHX_BLOCK) {
    const double *pc = part + c * NV;
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = red_slot<MINM, MAXM>(i, acc[i], pc[i]);
  }
  const double r = block_reduce_slots<NV, MINM, MAXM>(acc, lds);
  if (threadIdx.x < NV) res[threadIdx.x] = r;
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Split 64-bit fixed-point accumulation of per-slice sums.
// A value u with |u| <= B is scaled by 2^e, e = 60 - ilogb(B), so
// |q| = |rint(u*2^e)| < 2^61; q = hi*2^32 + lo with lo in [0, 2^32).
// Sums of hi (int64) and lo (uint64) are exact for up to 2^31 particles per
// slice, so the per-slice sum is independent of summation order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int fx_exponent(double bound) {
  if (!(bound > 0.0) || !isfinite(bound)) return 0;
  return 60 - ilogb(bound);
}

__device__ __forceinline__ void fx_split(double u, int e, unsigned long long &hi,
                                         unsigned long long &lo) {
  const long long q = llrint(ldexp(u, e));
  hi = (unsigned long long)(q >> 32);
  lo = (unsigned long long)(q & 0xffffffffLL);
}

__device__ __forceinline__ double fx_value(unsigned long long hi_sum, unsigned long long lo_sum,
                                           int e) {
  const long long h = (long long)hi_sum + (long long)(lo_sum >> 32);
  const double top = ldexp((double)h, 32);
  const double s = top + (double)(lo_sum & 0xffffffffULL);
  return ldexp(s, -e);
}

// ---------------------------------------------------------------------------
// Cuts of UniformBinSlicer (particles/slicing.py:395-413, 599-615) from the
// reduced z sums, and z_bins as np.arange(tail, head + 1e-7*width, width):
// buffer[0] = start, buffer[1] = start + step, delta = buffer[1] - start,
// buffer[i] = start + i*delta (NumPy's arange fill for float64).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void compute_cuts(const phx_slicer &cfg, int64_t n, double sum_z,
                                             double sum_dz, double sum_dz2, double zmin,
                                             double zmax, phx_slice_state &st) {
  double tail, head, mean_z = 0.0, sigma_z = 0.0;
  if (cfg.cuts_mode == PHX_CUTS_FIXED) {
    tail = cfg.z_cut_tail;
    head = cfg.z_cut_head;
  } else if (cfg.cuts_mode == PHX_CUTS_NSIGMA) {
    mean_z = sum_z / (double)n;
    double cov = 0.0;
    if (n >= 2) cov = (sum_dz2 - sum_dz * sum_dz / (double)n) / (double)(n - 1);
    sigma_z = sqrt(cov);
    tail = mean_z - cfg.n_sigma_z * sigma_z;
    head = mean_z + cfg.n_sigma_z * sigma_z;
  } else {
    tail = zmin;
    head = zmax;
    head = head + fabs(head) * 1e-15;
  }
  st.tail = tail;
  st.head = head;
  st.width = (head - tail) / (double)cfg.n_slices;
  st.mean_z = mean_z;
  st.sigma_z = sigma_z;
  st.n = n;
  st.n_slices = cfg.n_slices;
  st.valid = 1;
}

__device__ __forceinline__ double arange_value(double start, double step, int i) {
  if (i == 0) return start;
  const double second = start + step;
  if (i == 1) return second;
  const double delta = second - start;
  return start + (double)i * delta;
}
