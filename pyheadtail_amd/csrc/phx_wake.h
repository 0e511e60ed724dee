// Wake-field device code shared by the standalone calls and the engine.
//   wake functions     impedances/wakes.py:450-493 (resonator), :611-624 (resistive wall)
//   WakeKick engine    impedances/wake_kicks.py:66-139
//   WakeField history  impedances/wakes.py:116-145
#pragma once

#include "phx_device.h"

#define WK_TILE 32                   // target slices per block
#define WK_GROUPS (PHX_BLOCK / WK_TILE)

// Evaluation of the wake function with the reference's operation order.
// rw_pre[c] is the resistive-wall prefix ((((Y*(c-2))/2)*sqrt(beta))*L/pi)/b^3
// for sign(dt+|dt_min|) - 1 = c - 2 (precomputed per source turn; the chain
// only takes three values).
__device__ __forceinline__ double wake_eval(const phx_wake_kick &w, double dt,
                                            const double *rw_pre) {
  if (w.fn == PHX_WAKE_RES_TRANSVERSE) {
    const double d = dt > 0.0 ? 0.0 : dt;  // dt.clip(max=0)
    const double ea = w.prefix * exp(w.alpha * d);
    if (w.Q > 0.5) return ea * phx_sin(w.omegabar * d);
    if (w.Q == 0.5) return ea * d;
    return ea * sinh(w.omegabar * d);
  }
  if (w.fn == PHX_WAKE_RES_LONGITUDINAL) {
    const double sg = dt > 0.0 ? 1.0 : (dt < 0.0 ? -1.0 : (dt == 0.0 ? 0.0 : dt));
    const double t = ((-(sg - 1.0) * w.prefix) * w.alpha) * exp(w.alpha * dt);
    if (w.Q > 0.5) {
      double sn, cs;
      phx_sincos(w.omegabar * dt, &sn, &cs);
      return t * (cs + w.alpha_over_omegabar * sn);
    }
    if (w.Q == 0.5) return t * (1.0 + w.alpha * dt);
    return t * (cosh(w.omegabar * dt) + w.alpha_over_omegabar * sinh(w.omegabar * dt));
  }
  // resistive wall
  const double a = dt + w.dt_min_abs;
  double pre;
  if (a > 0.0) pre = rw_pre[2];
  else if (a < 0.0) pre = rw_pre[0];
  else if (a == 0.0) pre = rw_pre[1];
  else pre = a;  // NaN propagates like np.sign(nan)
  const double mdt = -w.dt_min_abs;
  const double dtc = dt > mdt ? mdt : dt;  // dt.clip(max=-|dt_min|)
  return (pre * sqrt(w.k_sigma / dtc)) * w.sqrt_z0c;
}

__device__ __forceinline__ void rw_prefixes(const phx_wake_kick &w, double beta_src,
                                            double *out3) {
  const double sb = sqrt(beta_src);
  for (int c = 0; c < 3; ++c) {
    double t = w.yokoya * (double)(c - 2);
    t = t / 2.0;
    t = t * sb;
    t = t * w.L;
    t = t / PHX_PI;
    t = t / w.b3;
    out3[c] = t;
  }
}

__device__ __forceinline__ int ring_slot(int head, int age_index, int capacity) {
  int s = (head - age_index) % capacity;
  return s < 0 ? s + capacity : s;
}

// Push one SliceSet into the history (called by one block).  Ages of the
// stored entries grow by age_inc first, then the new entry is added with
// age 0 (WakeField.track, impedances/wakes.py:134-142).
__device__ __forceinline__ void hist_push_device(phx_wake_hist *h, const double *times,
                                                 const int32_t *counts, const double *mean_x,
                                                 const double *mean_y, double age_inc,
                                                 double beta) {
  __shared__ int sh_head;
  const int cap = h->capacity, S = h->n_slices;
  if (threadIdx.x == 0) {
    const int nf = h->n_filled, head = h->head;
    for (int i = 0; i < nf; ++i) {
      const int sl = ring_slot(head, i, cap);
      h->age[sl] = h->age[sl] + age_inc;
    }
    const int nh = (head + 1) % cap;
    h->head = nh;
    h->n_filled = nf + 1 < cap ? nf + 1 : cap;
    h->age[nh] = 0.0;
    h->beta[nh] = beta;
    sh_head = nh;
  }
  lds_sync();
  const int nh = sh_head;
  double *tt = h->times + (size_t)nh * S;
  double *mm = h->mom + (size_t)nh * 3 * S;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const double c = (double)counts[s];
    tt[s] = times[s];
    mm[s] = c;
    mm[S + s] = mean_x ? c * mean_x[s] : 0.0;
    mm[2 * S + s] = mean_y ? c * mean_y[s] : 0.0;
  }
  lds_sync();
}

// One block = (target tile, kick, turn).  With one history slot the block
// writes the kick directly, wake_factor * (0 + conv); otherwise it writes the
// convolution of its turn into part[(k*T + t)*S + j] and wake_conv_reduce
// sums the turns in order (WakeKick._accumulate_source_signal,
// wake_kicks.py:118-139).
struct WakeKicks {
  phx_wake_kick k[PHX_MAX_KICKS];
};

__device__ __forceinline__ void wake_conv_block(const phx_wake_hist *h, const WakeKicks &kk,
                                                int S, double *part, int T, double *out) {
  // LDS: [WK_TILE + S - 1] wake values (convolution form) or [S] source
  // times (dot-product form), then [S] source moments
  extern __shared__ double vbuf[];
  __shared__ double red[PHX_BLOCK];
  __shared__ double rwp[3];
  const int tile = blockIdx.x, ki = blockIdx.y, t = blockIdx.z;
  const phx_wake_kick &w = kk.k[ki];
  const int nf = h->n_filled, head = h->head, cap = h->capacity;
  const int n_turns = nf < w.n_turns_wake ? nf : w.n_turns_wake;
  if (t >= n_turns) return;
  const int j0 = tile * WK_TILE;
  const int jj = threadIdx.x % WK_TILE, g = threadIdx.x / WK_TILE;
  const int j = j0 + jj;
  const int slot = ring_slot(head, t, cap);
  const double age = h->age[slot];
  const double *tsrc = h->times + (size_t)slot * S;
  const double *ttgt = h->times + (size_t)head * S;
  const double *a = h->mom + ((size_t)slot * 3 + w.moment) * S;
  double *abuf = vbuf + (WK_TILE + S - 1);
  for (int i = threadIdx.x; i < S; i += blockDim.x) abuf[i] = a[i];
  if (threadIdx.x == 0 && w.fn == PHX_WAKE_RESISTIVE_WALL) rw_prefixes(w, h->beta[slot], rwp);
  lds_sync();
  double acc = 0.0;
  if (w.dot_product) {
    // dt[i, j] = t_target[j] - (t_source[i] + age)     (wake_kicks.py:81-85)
    for (int i = threadIdx.x; i < S; i += blockDim.x) vbuf[i] = tsrc[i] + age;
    lds_sync();
    if (j < S) {
      const double tj = ttgt[j];
      for (int i = g; i < S; i += WK_GROUPS) acc = acc + abuf[i] * wake_eval(w, tj - vbuf[i], rwp);
    }
  } else {
    // dt = concat(t_tgt - s_src[-1], (t_tgt - s_src[0])[1:]) and
    // np.convolve(moments, W, 'valid')[j] = sum_m a[m] W[j + S - 1 - m]
    // (wake_kicks.py:109-116)
    const double s_last = tsrc[S - 1] + age, s_first = tsrc[0] + age;
    const int p0 = j0, p1 = min(j0 + WK_TILE - 1 + S - 1, 2 * S - 2);
    for (int p = p0 + threadIdx.x; p <= p1; p += blockDim.x) {
      const double dt = p <= S - 1 ? ttgt[p] - s_last : ttgt[p - S + 1] - s_first;
      vbuf[p - p0] = wake_eval(w, dt, rwp);
    }
    lds_sync();
    if (j < S) {
      for (int m = g; m < S; m += WK_GROUPS) acc = acc + abuf[m] * vbuf[j + S - 1 - m - p0];
    }
  }
  red[threadIdx.x] = acc;
  lds_sync();
  if (g == 0 && j < S) {
    double r = red[jj];
    for (int q = 1; q < WK_GROUPS; ++q) r = r + red[q * WK_TILE + jj];
    if (T == 1) out[(size_t)ki * S + j] = w.wake_factor * (0.0 + r);
    else part[((size_t)ki * T + t) * S + j] = r;
  }
}

// Sum of the per-turn convolutions in turn order, times the wake factor.
__device__ __forceinline__ void wake_conv_reduce(const phx_wake_hist *h, const WakeKicks &kk,
                                                 int n_kicks, int S, const double *part, int T,
                                                 double *out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_kicks * S) return;
  const int ki = idx / S, j = idx % S;
  const phx_wake_kick &w = kk.k[ki];
  const int nf = h->n_filled;
  const int n_turns = nf < w.n_turns_wake ? nf : w.n_turns_wake;
  double acc = 0.0;
  for (int q = 0; q < n_turns; ++q) acc = acc + part[((size_t)ki * T + q) * S + j];
  out[(size_t)ki * S + j] = w.wake_factor * acc;
}

// Dynamic LDS of wake_conv_block; the static part is PHX_BLOCK + 3 doubles.
inline size_t wake_conv_lds_bytes(int S) { return (size_t)(WK_TILE + 2 * S - 1) * sizeof(double); }
#define WK_STATIC_LDS ((PHX_BLOCK + 3) * sizeof(double))
