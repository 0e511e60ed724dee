// Device-side building blocks shared by the standalone element kernels and
// the fused turn engine.  Everything here is FP64 and is compiled with
// -ffp-contract=off so that each arithmetic operation is one IEEE-754
// rounding, in the order the reference's NumPy expressions evaluate them.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/phx.h"
#include "phx_math.h"

#define PHX_BLOCK 256
#define PHX_WAVE 64
#ifndef PHX_PPT
#define PHX_PPT 2
#endif
#define PHX_CHUNK (PHX_BLOCK * PHX_PPT)
// Reducing particle passes run a persistent grid of G blocks: block b
// handles the particle groups b, b + G, b + 2G, ... of PHX_CHUNK particles
// (thread t: particles g*PHX_CHUNK + k*PHX_BLOCK + t, k < PHX_PPT), so block
// b's partial covers a fixed set of particles in a fixed order for a given G.
// The standalone reductions use G = PHX_STANDALONE_GRID (three blocks on
// each of MI355X's 256 CUs); the engine's pass uses the grid that stays
// resident for its kernel, at most PHX_MAX_GRID.

__host__ __device__ inline int64_t phx_n_groups(int64_t n) {
  return n <= 0 ? 0 : (n + PHX_CHUNK - 1) / PHX_CHUNK;
}
__host__ __device__ inline int phx_grid_for(int64_t n, int64_t grid) {
  const int64_t g = phx_n_groups(n);
  return (int)(g < grid ? g : grid);
}
__host__ __device__ inline int phx_pass_grid(int64_t n) {
  return phx_grid_for(n, PHX_STANDALONE_GRID);
}
// NumPy's 2.*np.pi as a double
#define PHX_TWO_PI 6.283185307179586
#define PHX_PI 3.141592653589793

struct P6 {
  double x, xp, y, yp, z, dp;
};

// Workgroup barrier for data exchanged through LDS only.  __syncthreads()
// also waits for every outstanding global load and store of the wave (its
// release fence covers global memory); this one waits for LDS operations
// only, so particle stores stay in flight across it.  Every barrier of this
// library protects LDS data only.  No kernel here leaves a global load in
// flight across it: the slice histogram's fold of the cuts, with the first
// particle loads issued ahead of it, gave wrong maxima in some blocks on
// MI355X.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

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

// a / b, given r = 1.0 / b (correctly rounded): q0 = a r, the residual
// a - b q0 (exact, one FMA) and q0 + (a - b q0) r rounded once, which is the
// correctly rounded quotient when nothing overflows or underflows
// (Markstein's theorem on division by a correctly rounded reciprocal) --
// the same bits as the division, in three operations instead of the
// division's scaled Newton sequence.  Quotients outside [2^-900, 2^900]
// (zero, infinity, NaN, extreme magnitudes) are recomputed by the
// division itself.
__device__ __forceinline__ double div_rcp(double a, double b, double r) {
  const double q0 = a * r;
  const double e = fma(-q0, b, a);
  const double q = fma(e, r, q0);
  const double m = fabs(q);
  if (__builtin_expect(!(m >= 0x1p-900 && m <= 0x1p900), 0)) return a / b;
  return q;
}

// slice_floor with inv_width = 1.0 / width (the same value).
__device__ __forceinline__ double slice_floor_r(double z, double tail, double width,
                                                double inv_width) {
  return floor(div_rcp(z - tail, width, inv_width));
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
__device__ __forceinline__ void detune_amplitude(const phx_detuner &d, const P6 &p, double &qx,
                                                 double &qy) {
  // Jx = 1/2*((1+a^2)/b*x**2 + 2*a*x*xp + b*xp**2)
  const double jx = 0.5 * ((d.gx * (p.x * p.x) + (d.ax2 * p.x) * p.xp) + d.bx * (p.xp * p.xp));
  const double jy = 0.5 * ((d.gy * (p.y * p.y) + (d.ay2 * p.y) * p.yp) + d.by * (p.yp * p.yp));
  qx = div_rcp(d.dapp_x * jx + d.dapp_xy * jy, d.p0, d.inv_p0);
  qy = div_rcp(d.dapp_y * jy + d.dapp_yx * jx, d.p0, d.inv_p0);
}

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
    detune_amplitude(d, p, qx, qy);
  }
}

// Smooth segments (PHX_DETS_SMOOTH) have alpha = 0, and the host sets the
// flag only with ax2 = ay2 = 0 and gx, gy > 0: then (ax2 x) xp is a zero
// added to gx x^2 >= +0, which leaves it unchanged, and the term is dropped
// (the same bits for finite coordinates).
__device__ __forceinline__ void detune_amplitude_smooth(const phx_detuner &d, const P6 &p,
                                                        double &qx, double &qy) {
  const double jx = 0.5 * (d.gx * (p.x * p.x) + d.bx * (p.xp * p.xp));
  const double jy = 0.5 * (d.gy * (p.y * p.y) + d.by * (p.yp * p.yp));
  qx = div_rcp(d.dapp_x * jx + d.dapp_xy * jy, d.p0, d.inv_p0);
  qy = div_rcp(d.dapp_y * jy + d.dapp_yx * jx, d.p0, d.inv_p0);
}

// The detuner sets with straight-line code (PHX_DETS_*): first-order
// chromaticity (order-1 Horner: c0 * dp), then amplitude detuning; the
// arithmetic of detune() for those sets.
template <int DS>
__device__ __forceinline__ void detune_set(const phx_tseg &t, const P6 &p, double &dx,
                                           double &dy) {
  dx = t.dQ_x;
  dy = t.dQ_y;
  const phx_detuner &c = t.det[0];
  dx = dx + c.cx[0] * p.dp;
  dy = dy + c.cy[0] * p.dp;
  if constexpr ((DS & ~PHX_DETS_SMOOTH) == PHX_DETS_C1_AMP) {
    double qx, qy;
    if constexpr ((DS & PHX_DETS_SMOOTH) != 0) detune_amplitude_smooth(t.det[1], p, qx, qy);
    else detune_amplitude(t.det[1], p, qx, qy);
    dx = dx + qx;
    dy = dy + qy;
  }
}

// The element functions act on N particles at once: the element's constants
// are read once and the N independent particles interleave (instruction-level
// parallelism for the long FP64 chains).  The arithmetic of each particle is
// the same for every N.
template <int N, int DS = PHX_DETS_GENERIC>
__device__ __forceinline__ void apply_tseg_n(P6 (&p)[N], const phx_tseg &t) {
  double M00[N], M01[N], M10[N], M11[N], M22[N], M23[N], M32[N], M33[N];
  constexpr bool SMOOTH = (DS & PHX_DETS_SMOOTH) != 0 && (DS & ~PHX_DETS_SMOOTH) >= PHX_DETS_C1;
  const int nd = DS == PHX_DETS_GENERIC ? t.n_detuners : (DS == PHX_DETS_NONE ? 0 : 1);
  if (nd == 0) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      M00[k] = t.M00; M01[k] = t.M01; M10[k] = t.M10; M11[k] = t.M11;
      M22[k] = t.M22; M23[k] = t.M23; M32[k] = t.M32; M33[k] = t.M33;
    }
  } else {
    double dphi_x[N], dphi_y[N];
    if constexpr (DS == PHX_DETS_GENERIC) {
#pragma unroll
      for (int k = 0; k < N; ++k) {
        dphi_x[k] = t.dQ_x;
        dphi_y[k] = t.dQ_y;
      }
      for (int i = 0; i < nd; ++i) {
        const phx_detuner &d = t.det[i];
#pragma unroll
        for (int k = 0; k < N; ++k) {
          double qx, qy;
          detune(d, p[k], qx, qy);
          dphi_x[k] = dphi_x[k] + qx;
          dphi_y[k] = dphi_y[k] + qy;
        }
      }
    } else {
#pragma unroll
      for (int k = 0; k < N; ++k) detune_set<DS>(t, p[k], dphi_x[k], dphi_y[k]);
    }
    const double I00 = t.I00, I01 = t.I01, I10 = t.I10, I11 = t.I11;
    const double I22 = t.I22, I23 = t.I23, I32 = t.I32, I33 = t.I33;
    const double J00 = t.J00, J01 = t.J01, J10 = t.J10, J11 = t.J11;
    const double J22 = t.J22, J23 = t.J23, J32 = t.J32, J33 = t.J33;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const double ax = dphi_x[k] * PHX_TWO_PI;
      const double ay = dphi_y[k] * PHX_TWO_PI;
      double sx, cx, sy, cy;
      phx_sincos(ax, &sx, &cx);
      phx_sincos(ay, &sy, &cy);
      if constexpr (SMOOTH) {
        // 1 * cos + 0 * sin and 0 * cos + J * sin, each rounded once: the
        // same values
        M00[k] = cx;
        M01[k] = J01 * sx;
        M10[k] = J10 * sx;
        M11[k] = cx;
        M22[k] = cy;
        M23[k] = J23 * sy;
        M32[k] = J32 * sy;
        M33[k] = cy;
        continue;
      }
      M00[k] = I00 * cx + J00 * sx;
      M01[k] = I01 * cx + J01 * sx;
      M10[k] = I10 * cx + J10 * sx;
      M11[k] = I11 * cx + J11 * sx;
      M22[k] = I22 * cy + J22 * sy;
      M23[k] = I23 * cy + J23 * sy;
      M32[k] = I32 * cy + J32 * sy;
      M33[k] = I33 * cy + J33 * sy;
    }
  }
  const bool disp = !SMOOTH && t.with_dispersion != 0;
  const double dx0 = t.D_x_s0, dy0 = t.D_y_s0, dx1 = t.D_x_s1, dy1 = t.D_y_s1;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    P6 &q = p[k];
    if (disp) {
      q.x = q.x - dx0 * q.dp;
      q.y = q.y - dy0 * q.dp;
    }
    const double x = q.x, xp = q.xp, y = q.y, yp = q.yp;
    q.x = M00[k] * x + M01[k] * xp;
    q.xp = M10[k] * x + M11[k] * xp;
    q.y = M22[k] * y + M23[k] * yp;
    q.yp = M32[k] * y + M33[k] * yp;
    if (disp) {
      q.x = q.x + dx1 * q.dp;
      q.y = q.y + dy1 * q.dp;
    }
  }
}

// ---------------------------------------------------------------------------
// Longitudinal maps (trackers/longitudinal_tracking.py).
// ---------------------------------------------------------------------------
// STAT: a stationary RF sequence (PHX_SHAPE_STATIONARY): beta_ratio = 1,
// phi_offset = phi_lock = p_increment = +0 and no dispersion, checked by the
// host, so that 1 * z = z, (a + 0) + 0 = a + 0 and b - 0 = b drop those
// terms with the same bits.
template <int N, bool STAT = false>
__device__ __forceinline__ void drift_n(P6 (&p)[N], const phx_long_op &o) {
  // Drift.track: z = beta_ratio*z - eta(dp)*dp*length   (:107-111)
  const double eta0 = o.eta0, len = o.length;
  if constexpr (STAT) {
#pragma unroll
    for (int k = 0; k < N; ++k) p[k].z = p[k].z - (eta0 * p[k].dp) * len;
  } else {
    const double br = o.beta_ratio;
#pragma unroll
    for (int k = 0; k < N; ++k) p[k].z = br * p[k].z - (eta0 * p[k].dp) * len;
  }
}

template <int N, bool STAT = false>
__device__ __forceinline__ void rf_kick_n(P6 (&p)[N], const phx_long_op &o) {
  // Kick.track_with_dispersion / track_without_dispersion (:185-203)
  if constexpr (STAT) {
    const double h = o.harmonic, tp = o.two_pi, C = o.circumference;
    const double p0b = o.p0_before, p0a = o.p0_after, amp = o.amplitude;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      P6 &q = p[k];
      const double phi = h * div_rcp(tp * q.z, C, o.inv_circumference) + 0.0;
      const double delta_p = q.dp * p0b;
      q.dp = div_rcp(delta_p + amp * phx_sin(phi), p0a, o.inv_p0_after);
    }
    return;
  }
  const bool disp = o.with_dispersion != 0;
  const double Dx = o.D_x, Dy = o.D_y, h = o.harmonic, tp = o.two_pi, C = o.circumference;
  const double off = o.phi_offset, lock = o.phi_lock, p0b = o.p0_before, p0a = o.p0_after;
  const double amp = o.amplitude, pinc = o.p_increment;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    P6 &q = p[k];
    if (disp) {
      q.x = q.x - Dx * q.dp;
      q.y = q.y - Dy * q.dp;
    }
    const double phi = (h * div_rcp(tp * q.z, C, o.inv_circumference) + off) + lock;
    double delta_p = q.dp * p0b;
    delta_p = delta_p + (amp * phx_sin(phi) - pinc);
    q.dp = div_rcp(delta_p, p0a, o.inv_p0_after);
    if (disp) {
      q.x = q.x + Dx * q.dp;
      q.y = q.y + Dy * q.dp;
    }
  }
}

// A sequence is RFSystems' [drift, kicks..., drift] (:362-369) or a single
// Drift or Kick; the leading and trailing drifts are applied outside the
// loop over the inner operations, which keeps the kick's registers apart
// from the drift's.
template <int N>
__device__ __forceinline__ void shrink_n(P6 (&p)[N], const phx_long_seq &s) {
  // RFSystems._shrink_transverse_emittance (:556-565)
  if (s.shrink) {
    const double f = s.shrink_factor;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      p[k].x = p[k].x * f;
      p[k].xp = p[k].xp * f;
      p[k].y = p[k].y * f;
      p[k].yp = p[k].yp * f;
    }
  }
}

template <int N>
__device__ __forceinline__ void apply_long_seq_n(P6 (&p)[N], const phx_long_seq &s) {
  const int n_ops = s.n_ops;
  const bool lead = n_ops > 0 && s.op[0].type == PHX_LONG_DRIFT;
  const bool trail = n_ops > 1 && s.op[n_ops - 1].type == PHX_LONG_DRIFT;
  if (lead) drift_n<N>(p, s.op[0]);
  const int i1 = trail ? n_ops - 1 : n_ops;
  for (int i = lead ? 1 : 0; i < i1; ++i) {
    if (s.op[i].type == PHX_LONG_DRIFT) drift_n<N>(p, s.op[i]);
    else rf_kick_n<N>(p, s.op[i]);
  }
  if (trail) drift_n<N>(p, s.op[n_ops - 1]);
  shrink_n<N>(p, s);
}

// RFSystems' sequence with NH harmonics: op[0] drift, op[1..NH] kicks,
// op[NH+1] drift.
template <int NH, int N, bool STAT = false>
__device__ __forceinline__ void apply_rf_seq_n(P6 (&p)[N], const phx_long_seq &s) {
  drift_n<N, STAT>(p, s.op[0]);
#pragma unroll
  for (int i = 1; i <= NH; ++i) rf_kick_n<N, STAT>(p, s.op[i]);
  drift_n<N, STAT>(p, s.op[NH + 1]);
  shrink_n<N>(p, s);
}

template <int N>
__device__ __forceinline__ void apply_linmap_n(P6 (&p)[N], const phx_linmap &m) {
  // LinearMap.track_with_dispersion -> track_without_dispersion (:731-759)
  const double Dx = m.D_x, Dy = m.D_y, cq = m.cosQ, sq = m.sinQ, lf = m.longfac;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    P6 &q = p[k];
    q.x = q.x - Dx * q.dp;
    q.y = q.y - Dy * q.dp;
    const double z0 = q.z, dp0 = q.dp;
    q.z = z0 * cq - (lf * dp0) * sq;
    q.dp = dp0 * cq + (z0 / lf) * sq;
    q.x = q.x + Dx * q.dp;
    q.y = q.y + Dy * q.dp;
  }
}

template <int N>
__device__ __forceinline__ void apply_program_n(P6 (&p)[N], const phx_turn_program &g) {
  for (int e = 0; e < g.n_el; ++e) {
    const int k = g.el_index[e];
    switch (g.el_type[e]) {
      case PHX_EL_TSEG: apply_tseg_n<N>(p, g.tseg[k]); break;
      case PHX_EL_LONG: apply_long_seq_n<N>(p, g.lseq[k]); break;
      default: apply_linmap_n<N>(p, g.lin[k]); break;
    }
  }
}

// A program of a known element order (PHX_SHAPE_*): the elements are
// applied in straight-line code, which the compiler schedules and allocates
// registers for far better than the interpreter's loop over element kinds.
template <int SHAPE, int N, int DS = PHX_DETS_GENERIC>
__device__ __forceinline__ void apply_shape_n(P6 (&p)[N], const phx_turn_program &g) {
  constexpr int SH = SHAPE & ~PHX_SHAPE_STATIONARY;
  constexpr bool ST = SH != SHAPE;
  if constexpr (SH == PHX_SHAPE_RF1_T) {
    apply_rf_seq_n<1, N, ST>(p, g.lseq[0]);
    apply_tseg_n<N, DS>(p, g.tseg[0]);
  } else if constexpr (SH == PHX_SHAPE_RF2_T) {
    apply_rf_seq_n<2, N, ST>(p, g.lseq[0]);
    apply_tseg_n<N, DS>(p, g.tseg[0]);
  } else if constexpr (SH == PHX_SHAPE_T_RF1) {
    apply_tseg_n<N, DS>(p, g.tseg[0]);
    apply_rf_seq_n<1, N, ST>(p, g.lseq[0]);
  } else if constexpr (SH == PHX_SHAPE_T_RF2) {
    apply_tseg_n<N, DS>(p, g.tseg[0]);
    apply_rf_seq_n<2, N, ST>(p, g.lseq[0]);
  } else if constexpr (SH == PHX_SHAPE_LIN_T) {
    apply_linmap_n<N>(p, g.lin[0]);
    apply_tseg_n<N, DS>(p, g.tseg[0]);
  } else if constexpr (SH == PHX_SHAPE_T_LIN) {
    apply_tseg_n<N, DS>(p, g.tseg[0]);
    apply_linmap_n<N>(p, g.lin[0]);
  } else if constexpr (SH == PHX_SHAPE_T) {
    apply_tseg_n<N, DS>(p, g.tseg[0]);
  } else if constexpr (SH == PHX_SHAPE_NONE) {
  } else {
    apply_program_n<N>(p, g);
  }
}

// z and dp after a program: only the longitudinal elements change them
// (transverse segments act on x, xp, y, yp), so applying those alone gives
// the same z and dp bit for bit.
template <int SHAPE>
__device__ __forceinline__ void apply_shape_longitudinal(P6 &q, const phx_turn_program &g) {
  constexpr int SH = SHAPE & ~PHX_SHAPE_STATIONARY;
  constexpr bool ST = SH != SHAPE;
  P6(&p)[1] = *reinterpret_cast<P6(*)[1]>(&q);
  if constexpr (SH == PHX_SHAPE_RF1_T || SH == PHX_SHAPE_T_RF1) {
    apply_rf_seq_n<1, 1, ST>(p, g.lseq[0]);
  } else if constexpr (SH == PHX_SHAPE_RF2_T || SH == PHX_SHAPE_T_RF2) {
    apply_rf_seq_n<2, 1, ST>(p, g.lseq[0]);
  } else if constexpr (SH == PHX_SHAPE_LIN_T || SH == PHX_SHAPE_T_LIN) {
    apply_linmap_n<1>(p, g.lin[0]);
  } else if constexpr (SH == PHX_SHAPE_GENERIC) {
    for (int e = 0; e < g.n_el; ++e) {
      const int k = g.el_index[e];
      if (g.el_type[e] == PHX_EL_LONG) apply_long_seq_n<1>(p, g.lseq[k]);
      else if (g.el_type[e] == PHX_EL_LINMAP) apply_linmap_n<1>(p, g.lin[k]);
    }
  }
}

__device__ __forceinline__ P6 (&one(P6 &p))[1] { return *reinterpret_cast<P6(*)[1]>(&p); }
__device__ __forceinline__ void apply_tseg(P6 &p, const phx_tseg &t) { apply_tseg_n<1>(one(p), t); }
__device__ __forceinline__ void apply_long_seq(P6 &p, const phx_long_seq &s) {
  apply_long_seq_n<1>(one(p), s);
}
__device__ __forceinline__ void apply_linmap(P6 &p, const phx_linmap &m) {
  apply_linmap_n<1>(one(p), m);
}
__device__ __forceinline__ void apply_program(P6 &p, const phx_turn_program &g) {
  apply_program_n<1>(one(p), g);
}

// ---------------------------------------------------------------------------
// Wake kick application (impedances/wake_kicks.py:160-361), kicks in order.
// ---------------------------------------------------------------------------
// Kick k of slice idx is kicks[k * n_slices + idx] (the [kick][slice]
// layout of the C ABI); the table is small and stays in the caches.
__device__ __forceinline__ void apply_wake_kicks(P6 &p, int32_t idx, int n_slices,
                                                 const phx_wake_kick *kd, int n_kicks,
                                                 const double *__restrict__ kicks) {
  if (idx < 0 || idx >= n_slices) return;
  for (int k = 0; k < n_kicks; ++k) {
    const phx_wake_kick &d = kd[k];
    double v = kicks[(size_t)k * n_slices + idx];
    if (d.multiply == 0) v = v * p.x;
    else if (d.multiply == 2) v = v * p.y;
    if (d.plane == PHX_PLANE_XP) p.xp = p.xp + v;
    else if (d.plane == PHX_PLANE_YP) p.yp = p.yp + v;
    else p.dp = p.dp + v;
  }
}

// The kick step for a known kick set (PHX_KICKS_*).
template <int KS>
__device__ __forceinline__ void apply_kick_set(P6 &p, int32_t idx, int n_slices,
                                               const phx_wake_kick *kd, int n_kicks,
                                               const double *__restrict__ kicks) {
  if constexpr (KS == PHX_KICKS_DXDY) {
    if (idx < 0 || idx >= n_slices) return;
    p.xp = p.xp + kicks[idx];
    p.yp = p.yp + kicks[(size_t)n_slices + idx];
  } else if constexpr (KS == PHX_KICKS_PLATES) {
    if (idx < 0 || idx >= n_slices) return;
    p.xp = p.xp + kicks[idx];
    p.xp = p.xp + kicks[(size_t)n_slices + idx] * p.x;
    p.yp = p.yp + kicks[(size_t)2 * n_slices + idx];
    p.yp = p.yp + kicks[(size_t)3 * n_slices + idx] * p.y;
  } else {
    apply_wake_kicks(p, idx, n_slices, kd, n_kicks, kicks);
  }
}

// ---------------------------------------------------------------------------
// Deterministic reductions.  Within a block: every thread stores its slot
// values in LDS; for each slot, 16 threads each combine a fixed stride of
// the PHX_BLOCK values (lanes q, q+16, q+32, ...), then one thread combines
// the 16 results in order.  Across chunks: each thread folds a fixed stride
// of chunks in increasing order, then the same block reduction.  For a given
// particle count the association order is fixed, so results are
// bit-reproducible, and a slot's result does not depend on which other
// slots are reduced with it.  (An LDS transpose instead of wave shuffles:
// the shuffle tree is a chain of dependent cross-lane permutes per slot.)
// The operation of slot i is min if bit i of MINM is set, max if bit i of
// MAXM is set, else sum.
// ---------------------------------------------------------------------------
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

#define BR_GROUP 16  // slots per LDS round
#define BR_PARTS 16  // partial results per slot

// LDS doubles needed by block_reduce_slots for NV slots.
#define BR_LDS(NV) (((NV) < BR_GROUP ? (NV) : BR_GROUP) * (PHX_BLOCK + BR_PARTS))

// Block-reduce NV slots held by the first PHX_BLOCK threads of the block
// (every thread of the block must call it; larger blocks contribute only
// their first PHX_BLOCK threads).  Thread t < NV gets the block result of
// slot t as the return value (other threads: 0).  Slot t's operation is that
// of bit OFF + t of the masks.  lds: __shared__ double[BR_LDS(NV)].
template <int NV, unsigned long long MINM, unsigned long long MAXM, int OFF = 0>
__device__ __forceinline__ double block_reduce_slots(const double (&v)[NV], double *lds) {
  constexpr int G = NV < BR_GROUP ? NV : BR_GROUP;
  constexpr int STRIDE = PHX_BLOCK / BR_PARTS;
  double *part = lds + G * PHX_BLOCK;
  const int t = threadIdx.x;
  double out = 0.0;
#pragma unroll
  for (int g0 = 0; g0 < NV; g0 += G) {
    if (t < PHX_BLOCK) {
#pragma unroll
      for (int i = 0; i < G; ++i)
        if (g0 + i < NV) lds[i * PHX_BLOCK + t] = v[g0 + i];
    }
    lds_sync();
    if (t < G * BR_PARTS) {
      const int i = t / BR_PARTS, q = t % BR_PARTS;
      if (g0 + i < NV) {
        const int k = OFF + g0 + i;
        const double *row = lds + i * PHX_BLOCK;
        double r = row[q];
#pragma unroll
        for (int m = 1; m < STRIDE; ++m) r = red_slot<MINM, MAXM>(k, r, row[m * BR_PARTS + q]);
        part[i * BR_PARTS + q] = r;
      }
    }
    lds_sync();
    if (t >= g0 && t < g0 + G && t < NV) {
      const int i = t - g0, k = OFF + t;
      double r = part[i * BR_PARTS];
#pragma unroll
      for (int q = 1; q < BR_PARTS; ++q) r = red_slot<MINM, MAXM>(k, r, part[i * BR_PARTS + q]);
      out = r;
    }
    lds_sync();
  }
  return out;
}

// Reduce per-chunk partials part[c*NV + i] over c < n_chunks into res[i]
// (__shared__, valid on all threads after return).  Called by one block.
// Fold of the per-block partials of a reducing pass, stored slot-major:
// slot i of block c is part[i * nc + c].  One wave folds one slot: lane l
// combines the partials l, l + 64, l + 128, ... in increasing order, then
// the 64 lane values combine in a fixed butterfly (exchange distance 32,
// 16, ..., 1; the operations are commutative, so every lane ends with the
// same value).  The order depends on nc only, not on the block size, so
// every fold of the same partials gives the same bits.  The waves of the
// block take slots i0 + w, i0 + w + W, ...; res[i] receives slot i.  Ends
// with a barrier.
template <unsigned long long MINM, unsigned long long MAXM>
__device__ __forceinline__ void fold_partials(const double *part, int64_t nc, int i0, int i1,
                                              double *res) {
  const int lane = threadIdx.x % PHX_WAVE, wv = threadIdx.x / PHX_WAVE;
  const int nw = blockDim.x / PHX_WAVE;
  for (int i = i0 + wv; i < i1; i += nw) {
    const double *row = part + (size_t)i * nc;
    const double id = red_ident<MINM, MAXM>(i);
    double acc = id;
    // loads in batches, so that they are in flight together; padding with
    // the identity leaves the combination order unchanged
    constexpr int B = 8;
    for (int64_t c0 = lane; c0 < nc; c0 += (int64_t)B * PHX_WAVE) {
      double v[B];
#pragma unroll
      for (int k = 0; k < B; ++k) {
        const int64_t c = c0 + (int64_t)k * PHX_WAVE;
        v[k] = c < nc ? row[c] : id;
      }
#pragma unroll
      for (int k = 0; k < B; ++k) acc = red_slot<MINM, MAXM>(i, acc, v[k]);
    }
#pragma unroll
    for (int off = PHX_WAVE / 2; off > 0; off >>= 1)
      acc = red_slot<MINM, MAXM>(i, acc, __shfl_xor(acc, off));
    if (lane == 0) res[i] = acc;
  }
  lds_sync();
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
