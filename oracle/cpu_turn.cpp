// Test infrastructure only: a multi-core C++ restatement of the reference's
// CPU path for the benchmark's one-turn map, used as the all-core CPU
// baseline and checked against the NumPy oracle (oracle/refpath.py), which
// the reference's own golden vectors pin.  Nothing in the product links or
// calls it.
//
// One turn, in the reference's order (machines/synchrotron.py:515-611 with
// RF_at='middle'; the wake field appended by the user):
//   RFSystems       Drift(C/2), Kick, Drift(C/2)
//                   (trackers/longitudinal_tracking.py:84-111, 114-203,
//                    362-369)
//   TransverseSegmentMap with chromaticity (first order) and amplitude
//                   detuning (trackers/transverse_tracking.py:138-196,
//                    trackers/detuners.py:64-98, 128-151)
//   WakeField       UniformBinSlicer(n_sigma_z) cuts from mean z and the
//                   shifted covariance (particles/slicing.py:395-413,
//                    599-615; cobra_functions/stats.pyx:17-53), per-slice
//                   mean x / y (stats.pyx:235-253), np.convolve 'valid' of
//                   n*mean with the resonator wake (impedances/
//                   wake_kicks.py:87-116, impedances/wakes.py:450-470),
//                   dipolar x / y kicks (wake_kicks.py:160-240).
// Every per-particle expression keeps the IEEE operation order of the
// reference's NumPy expressions (built with -ffp-contract=off).  Sums are
// per-thread partials combined in thread order, so the z moments differ
// from np.mean's pairwise sum and the sequential cov in the last bits only.
// The wake kick of turn t is applied at the start of turn t+1's particle
// loop (it needs only z and turn t's cuts, which nothing changes in
// between), and after the last turn by orc_flush.
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

extern "C" {

typedef struct {
  // RFSystems, one harmonic, no acceleration
  double eta0, drift_length;
  double amplitude, harmonic, circumference, phi_offset, phi_lock, p0;
  // transverse segment: M = I cos + J sin
  double dQ_x, dQ_y;
  double I00, I01, I10, I11, I22, I23, I32, I33;
  double J00, J01, J10, J11, J22, J23, J32, J33;
  double chroma_x, chroma_y;  // Q'_x dmu, Q'_y dmu (first order)
  double gx, ax2, bx, gy, ay2, by;  // (1+alpha^2)/beta, 2 alpha, beta
  double dapp_x, dapp_y, dapp_xy, dapp_yx, det_p0;
  int32_t has_chroma, has_amplitude;
  // wake field
  int32_t n_slices, pad;
  double n_sigma_z, beta_c, wake_factor;
  double res_prefix, res_alpha, res_omegabar;  // yokoya R w^2/(Q wbar), w/2Q, wbar
} orc_ring;

typedef struct {
  double tail, width;
  int32_t pending, pad;
  double *kick_x, *kick_y;  // [n_slices]
} orc_state;

int orc_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  int t = 0;
#pragma omp parallel
  {
#pragma omp single
    t = omp_get_num_threads();
  }
  return t;
}

}  // extern "C"

namespace {

inline double slice_floor(double z, double tail, double width) {
  return floor((z - tail) / width);
}

inline void kick_particle(double z, double &xp, double &yp, const orc_state &st, int S) {
  const double f = slice_floor(z, st.tail, st.width);
  if (f >= 0.0 && f < (double)S) {
    const int s = (int)f;
    xp = xp + st.kick_x[s];
    yp = yp + st.kick_y[s];
  }
}

inline void rf_and_transverse(const orc_ring &r, double &x, double &xp, double &y, double &yp,
                              double &z, double &dp) {
  // Drift: z = beta_ratio*z - eta*dp*length with beta_ratio = 1, eta = eta0
  z = 1.0 * z - (r.eta0 * dp) * r.drift_length;
  // Kick
  const double phi = (r.harmonic * ((6.283185307179586 * z) / r.circumference) + r.phi_offset) +
                     r.phi_lock;
  double delta_p = dp * r.p0;
  delta_p = delta_p + (r.amplitude * sin(phi) - 0.0);
  dp = delta_p / r.p0;
  z = 1.0 * z - (r.eta0 * dp) * r.drift_length;
  // transverse segment
  double dphi_x = r.dQ_x, dphi_y = r.dQ_y;
  if (r.has_chroma) {
    dphi_x = dphi_x + r.chroma_x * dp;
    dphi_y = dphi_y + r.chroma_y * dp;
  }
  if (r.has_amplitude) {
    const double jx = 0.5 * ((r.gx * (x * x) + (r.ax2 * x) * xp) + r.bx * (xp * xp));
    const double jy = 0.5 * ((r.gy * (y * y) + (r.ay2 * y) * yp) + r.by * (yp * yp));
    dphi_x = dphi_x + (r.dapp_x * jx + r.dapp_xy * jy) / r.det_p0;
    dphi_y = dphi_y + (r.dapp_y * jy + r.dapp_yx * jx) / r.det_p0;
  }
  dphi_x = dphi_x * 6.283185307179586;
  dphi_y = dphi_y * 6.283185307179586;
  const double sx = sin(dphi_x), cx = cos(dphi_x), sy = sin(dphi_y), cy = cos(dphi_y);
  const double M00 = r.I00 * cx + r.J00 * sx, M01 = r.I01 * cx + r.J01 * sx;
  const double M10 = r.I10 * cx + r.J10 * sx, M11 = r.I11 * cx + r.J11 * sx;
  const double M22 = r.I22 * cy + r.J22 * sy, M23 = r.I23 * cy + r.J23 * sy;
  const double M32 = r.I32 * cy + r.J32 * sy, M33 = r.I33 * cy + r.J33 * sy;
  const double x0 = x, xp0 = xp, y0 = y, yp0 = yp;
  x = M00 * x0 + M01 * xp0;
  xp = M10 * x0 + M11 * xp0;
  y = M22 * y0 + M23 * yp0;
  yp = M32 * y0 + M33 * yp0;
}

inline double arange_value(double start, double step, int i) {
  if (i == 0) return start;
  const double second = start + step;
  if (i == 1) return second;
  return start + (double)i * (second - start);
}

double resonator(const orc_ring &r, double dt) {
  dt = dt > 0.0 ? 0.0 : dt;  // clip(max=0)
  return (r.res_prefix * exp(r.res_alpha * dt)) * sin(r.res_omegabar * dt);
}

}  // namespace

extern "C" {

// Applies the pending wake kick of the last turn (if any).
void orc_flush(double *x, double *xp, double *y, double *yp, double *z, double *dp, int64_t n,
               const orc_ring *ring, orc_state *st) {
  (void)x; (void)y; (void)dp;
  if (!st->pending) return;
  const int S = ring->n_slices;
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n; ++j) kick_particle(z[j], xp[j], yp[j], *st, S);
  st->pending = 0;
}

// n_turns turns of [RFSystems, TransverseSegmentMap, WakeField].
// Returns 0, or -1 on bad arguments.
int orc_turns(double *x, double *xp, double *y, double *yp, double *z, double *dp, int64_t n,
              const orc_ring *ring, orc_state *st, int32_t n_turns) {
  if (n < 2 || !ring || !st || ring->n_slices < 1 || !st->kick_x || !st->kick_y) return -1;
  const orc_ring r = *ring;
  const int S = r.n_slices;
  const int nt = omp_get_max_threads();
  std::vector<double> zsum(nt), dzsum(nt), dz2sum(nt);
  std::vector<double> hx((size_t)nt * S), hy((size_t)nt * S);
  std::vector<int64_t> hc((size_t)nt * S);
  std::vector<double> times(S), mx(S), my(S), wk(2 * S - 1);
  for (int turn = 0; turn < n_turns; ++turn) {
    // the shift of the covariance is particle 0's z after this turn's map
    double q[6] = {x[0], xp[0], y[0], yp[0], z[0], dp[0]};
    if (st->pending) kick_particle(q[4], q[1], q[3], *st, S);
    rf_and_transverse(r, q[0], q[1], q[2], q[3], q[4], q[5]);
    const double shift = q[4];
    const int pend = st->pending;
#pragma omp parallel
    {
      const int t = omp_get_thread_num();
      double a = 0.0, b = 0.0, c = 0.0;
#pragma omp for schedule(static)
      for (int64_t j = 0; j < n; ++j) {
        double X = x[j], XP = xp[j], Y = y[j], YP = yp[j], Z = z[j], DP = dp[j];
        if (pend) kick_particle(Z, XP, YP, *st, S);
        rf_and_transverse(r, X, XP, Y, YP, Z, DP);
        x[j] = X; xp[j] = XP; y[j] = Y; yp[j] = YP; z[j] = Z; dp[j] = DP;
        const double d = Z - shift;
        a = a + Z;
        b = b + d;
        c = c + d * d;
      }
      zsum[t] = a;
      dzsum[t] = b;
      dz2sum[t] = c;
    }
    st->pending = 0;
    double sz = 0.0, sdz = 0.0, sdz2 = 0.0;
    for (int t = 0; t < nt; ++t) {
      sz += zsum[t];
      sdz += dzsum[t];
      sdz2 += dz2sum[t];
    }
    const double mean_z = sz / (double)n;
    const double sigma_z = sqrt((sdz2 - sdz * sdz / (double)n) / (double)(n - 1));
    const double tail = mean_z - r.n_sigma_z * sigma_z;
    const double head = mean_z + r.n_sigma_z * sigma_z;
    const double width = (head - tail) / (double)S;
#pragma omp parallel
    {
      const int t = omp_get_thread_num();
      double *px = &hx[(size_t)t * S];
      double *py = &hy[(size_t)t * S];
      int64_t *pc = &hc[(size_t)t * S];
      for (int s = 0; s < S; ++s) {
        px[s] = 0.0;
        py[s] = 0.0;
        pc[s] = 0;
      }
#pragma omp for schedule(static)
      for (int64_t j = 0; j < n; ++j) {
        const double f = slice_floor(z[j], tail, width);
        if (f >= 0.0 && f < (double)S) {
          const int s = (int)f;
          px[s] += x[j];
          py[s] += y[j];
          pc[s] += 1;
        }
      }
    }
    for (int s = 0; s < S; ++s) {
      double sx = 0.0, sy = 0.0;
      int64_t c = 0;
      for (int t = 0; t < nt; ++t) {
        sx += hx[(size_t)t * S + s];
        sy += hy[(size_t)t * S + s];
        c += hc[(size_t)t * S + s];
      }
      const double cnt = (double)c;
      const double mean_x = c ? sx / cnt : 0.0, mean_y = c ? sy / cnt : 0.0;
      mx[s] = cnt * mean_x;
      my[s] = cnt * mean_y;
      const double zl = arange_value(tail, width, s), zr = arange_value(tail, width, s + 1);
      times[s] = (zl + 0.5 * (zr - zl)) / r.beta_c;
    }
    // dt = concat(t - t[-1], (t - t[0])[1:]); np.convolve(m, W(dt), 'valid')
    for (int k = 0; k < S; ++k) wk[k] = resonator(r, times[k] - times[S - 1]);
    for (int k = 1; k < S; ++k) wk[S - 1 + k] = resonator(r, times[k] - times[0]);
    for (int k = 0; k < S; ++k) {
      double ax = 0.0, ay = 0.0;
      for (int j = 0; j < S; ++j) {
        const double w = wk[k - j + S - 1];
        ax = ax + mx[j] * w;
        ay = ay + my[j] * w;
      }
      st->kick_x[k] = r.wake_factor * ax;
      st->kick_y[k] = r.wake_factor * ay;
    }
    st->tail = tail;
    st->width = width;
    st->pending = 1;
  }
  return 0;
}

}  // extern "C"
