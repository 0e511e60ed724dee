// Wake-field device code shared by the standalone calls and the engine.
//   wake functions     impedances/wakes.py:450-493 (resonator), :611-624 (resistive wall),
//                      :290-376 (wake tables)
//   WakeKick engine    impedances/wake_kicks.py:66-139
//   WakeField history  impedances/wakes.py:116-145
#pragma once

#include "phx_device.h"

#define WK_TILE 32                   // target slices per block
#define WK_THREADS 1024              // threads per block: at most one wake value each
#define WK_GROUPS (WK_THREADS / WK_TILE)

// Evaluation of the wake function with the reference's operation order, in
// two parts: the shape, which kicks of the same source share (they differ
// only in the Yokoya factor or the prefix that carries it), and the value of
// one kick from the shape.  wake_value(w, wake_shape(w, dt)) is the
// function's value bit for bit.
//   resonator, transverse:   W = (prefix * e^{a d}) * sin(wbar d)   (d = min(dt, 0))
//   resonator, longitudinal: W = (((-(sgn(dt)-1) * R) * a) * e^{a dt})
//                                * (cos(wbar dt) + a/wbar sin(wbar dt))
//   resistive wall:          W = (pre[c] * sqrt(k_sigma / dtc)) * sqrt(Z0 c)
// with rw_pre[c] the resistive-wall prefix ((((Y*(c-2))/2)*sqrt(beta))*L/pi)/b^3
// for sign(dt+|dt_min|) - 1 = c - 2 (per source turn; the chain only takes
// three values) and dtc = min(dt, -|dt_min|).
struct WakeShape {
  double f;  // e^{a d}; longitudinal: -(sgn(dt) - 1); resistive wall: class c, or NaN
  double g;  // sin / d / sinh; longitudinal: e^{a dt}; resistive wall: sqrt(k_sigma/dtc)
  double h;  // longitudinal: cos + a/wbar sin (cosh + a/wbar sinh, 1 + a dt)
};

// np.interp(x, t, w) for x inside [t[0], t[n-1]] (numpy's arr_interp: the
// node value where x hits a node, else slope * (x - t[j]) + w[j] with the
// other end's form if that gives NaN); left / right outside, NaN for NaN.
__device__ __forceinline__ double table_interp(const double *t, const double *w, int n, double x,
                                               double left, double right) {
  if (x != x) return x;
  if (x < t[0]) return left;
  if (x > t[n - 1]) return right;
  if (x == t[n - 1]) return w[n - 1];
  int lo = 0, hi = n - 1;  // t[lo] <= x < t[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (t[mid] <= x) lo = mid;
    else hi = mid;
  }
  if (t[lo] == x) return w[lo];
  const double slope = (w[lo + 1] - w[lo]) / (t[lo + 1] - t[lo]);
  double r = slope * (x - t[lo]) + w[lo];
  if (r != r) {
    r = slope * (x - t[lo + 1]) + w[lo + 1];
    if (r != r && w[lo] == w[lo + 1]) r = w[lo];
  }
  return r;
}

// WakeTable's wake functions (impedances/wakes.py:290-376), see PHX_TABLE_*.
__device__ __forceinline__ double table_wake(const phx_wake_kick &w, double dt) {
  const double *t = w.table_t, *v = w.table_w;
  const int n = w.table_n;
  switch (w.table_kind) {
    case PHX_TABLE_TRANS_UR: {
      const double d = dt > 0.0 ? 0.0 : dt;  // dt.clip(max=0)
      return table_interp(t, v, n, -d, NAN, NAN);
    }
    case PHX_TABLE_TRANS_LOWBETA: return table_interp(t, v, n, -dt, NAN, NAN);
    case PHX_TABLE_LONG_BL: {
      const double x = -dt;
      const double sg = x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : (x == 0.0 ? 0.0 : x));
      return (sg + 1.0) / 2.0 * table_interp(t, v, n, x, 0.0, 0.0);
    }
    default: return table_interp(t, v, n, -dt, 0.0, 0.0);
  }
}

__device__ __forceinline__ WakeShape wake_shape(const phx_wake_kick &w, double dt) {
  WakeShape b;
  b.h = 0.0;
  if (w.fn == PHX_WAKE_RES_TRANSVERSE) {
    const double d = dt > 0.0 ? 0.0 : dt;  // dt.clip(max=0)
    b.f = exp(w.alpha * d);
    if (w.Q > 0.5) b.g = phx_sin(w.omegabar * d);
    else if (w.Q == 0.5) b.g = d;
    else b.g = sinh(w.omegabar * d);
  } else if (w.fn == PHX_WAKE_RES_LONGITUDINAL) {
    const double sg = dt > 0.0 ? 1.0 : (dt < 0.0 ? -1.0 : (dt == 0.0 ? 0.0 : dt));
    b.f = -(sg - 1.0);
    b.g = exp(w.alpha * dt);
    if (w.Q > 0.5) {
      double sn, cs;
      phx_sincos(w.omegabar * dt, &sn, &cs);
      b.h = cs + w.alpha_over_omegabar * sn;
    } else if (w.Q == 0.5) {
      b.h = 1.0 + w.alpha * dt;
    } else {
      b.h = cosh(w.omegabar * dt) + w.alpha_over_omegabar * sinh(w.omegabar * dt);
    }
  } else if (w.fn == PHX_WAKE_TABLE) {
    b.f = table_wake(w, dt);
    b.g = 0.0;
  } else {
    const double a = dt + w.dt_min_abs;
    if (a > 0.0) b.f = 2.0;
    else if (a < 0.0) b.f = 0.0;
    else if (a == 0.0) b.f = 1.0;
    else b.f = a;  // NaN propagates like np.sign(nan)
    const double mdt = -w.dt_min_abs;
    const double dtc = dt > mdt ? mdt : dt;  // dt.clip(max=-|dt_min|)
    b.g = sqrt(w.k_sigma / dtc);
  }
  return b;
}

__device__ __forceinline__ double wake_value(const phx_wake_kick &w, const WakeShape &b,
                                             const double *rw_pre) {
  if (w.fn == PHX_WAKE_RES_TRANSVERSE) return (w.prefix * b.f) * b.g;
  if (w.fn == PHX_WAKE_RES_LONGITUDINAL) return (((b.f * w.prefix) * w.alpha) * b.g) * b.h;
  if (w.fn == PHX_WAKE_TABLE) return b.f;
  const double pre = b.f == 0.0 ? rw_pre[0] : (b.f == 1.0 ? rw_pre[1] : (b.f == 2.0 ? rw_pre[2] : b.f));
  return (pre * b.g) * w.sqrt_z0c;
}

__device__ __forceinline__ double wake_eval(const phx_wake_kick &w, double dt,
                                            const double *rw_pre) {
  return wake_value(w, wake_shape(w, dt), rw_pre);
}

// Kicks whose functions share the shape: same function, same parameters
// apart from the Yokoya factor, same memory and convolution form.  Every
// field is compared (no early exit), so that the loads of a scan over the
// kick table issue together rather than one after the other.
__device__ __forceinline__ bool same_shape(const phx_wake_kick &a, const phx_wake_kick &b) {
  const bool head = (a.fn == b.fn) & (a.n_turns_wake == b.n_turns_wake) &
                    (a.dot_product == b.dot_product) & (a.precision == b.precision);
  const bool res = (a.Q == b.Q) & (a.alpha == b.alpha) & (a.omegabar == b.omegabar);
  const bool lon = res & (a.alpha_over_omegabar == b.alpha_over_omegabar);
  const bool tab = (a.table_t == b.table_t) & (a.table_w == b.table_w) &
                   (a.table_n == b.table_n) & (a.table_kind == b.table_kind);
  const bool rw = (a.L == b.L) & (a.b3 == b.b3) & (a.dt_min_abs == b.dt_min_abs) &
                  (a.k_sigma == b.k_sigma) & (a.sqrt_z0c == b.sqrt_z0c);
  const bool body = a.fn == PHX_WAKE_RES_TRANSVERSE     ? res
                    : a.fn == PHX_WAKE_RES_LONGITUDINAL ? lon
                    : a.fn == PHX_WAKE_TABLE            ? tab
                                                        : rw;
  return head & body;
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
  const int cap = h->capacity, S = h->n_slices;
  const int nf = h->n_filled, head = h->head;
  const int nh = (head + 1) % cap;
  // one thread per stored entry; the slot of the new entry (ring full) is
  // set to age 0 below
  for (int i = threadIdx.x; i < nf; i += blockDim.x) {
    const int sl = ring_slot(head, i, cap);
    if (sl != nh) h->age[sl] = h->age[sl] + age_inc;
  }
  lds_sync();  // every thread has read the ring's state before thread 0 moves it
  if (threadIdx.x == 0) {
    h->head = nh;
    h->n_filled = nf + 1 < cap ? nf + 1 : cap;
    h->age[nh] = 0.0;
    h->beta[nh] = beta;
  }
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
//
// In the dot-product form (multi-turn memory), where every (source, target)
// pair needs its own wake value, the kicks that share the function's shape
// (same_shape, in index order) are served by one block, up to gmax of them:
// the block of the member at position 0, gmax, 2 gmax, ... evaluates the
// shape once per pair and the value of each member from it; the other
// members' blocks return at once.  Each kick's value and sum are those of a
// block of its own, bit for bit.
struct WakeKicks {
  phx_wake_kick k[PHX_MAX_KICKS];
};

// Copy of the first nk kicks into dst (LDS) by every thread of the block,
// then a barrier.  The blocks scan the table for the members of a shape; in
// global memory (the engine's turn programs) each step of such a scan waits
// for a load from L2, in LDS for an LDS read.
static_assert(sizeof(phx_wake_kick) % sizeof(uint64_t) == 0, "kick records are 8-byte words");
__device__ __forceinline__ void stage_kicks(const WakeKicks &src, int nk, WakeKicks &dst) {
  constexpr int words = (int)(sizeof(phx_wake_kick) / sizeof(uint64_t));
  const uint64_t *s = reinterpret_cast<const uint64_t *>(src.k);
  uint64_t *d = reinterpret_cast<uint64_t *>(dst.k);
  for (int i = threadIdx.x; i < nk * words; i += blockDim.x) d[i] = s[i];
  lds_sync();
}

// Resistive-wall kicks in the dot-product form have a kernel of their own
// (wake_rw_block); the general block serves every other kick.
__device__ __host__ __forceinline__ bool is_rw_dot(const phx_wake_kick &k) {
  return k.dot_product && k.fn == PHX_WAKE_RESISTIVE_WALL;
}

#define WK_GMAX 4  // kicks one block serves

// Rows of partial sums per kick.  The sources of the current turn, whose
// pairs straddle dt = -|dt_min| and are summed pair by pair at every
// precision, are split over rw_chunks(S) blocks per target tile (up to
// RW_CHUNK_MAX of them, each with a row of its own), so that this turn runs on
// as many compute units as the older ones; the split depends on S alone, so
// the sums do not depend on the memory depth T of the launch.  Row r of a
// resistive-wall kick in the dot-product form: r < nc, chunk r of turn 0;
// r >= nc, turn r - nc + 1.  Every other kick: row t, turn t (with T = 1 the
// general block writes the kick itself).  The reduction adds a kick's rows in
// order; a resistive-wall kick with a single row (S <= RW_CHUNK, T = 1) is
// written by its block.
#define RW_CHUNK 64
#define RW_CHUNK_MAX 16
__host__ __device__ inline int rw_chunks(int S) {
  const int c = (S + RW_CHUNK - 1) / RW_CHUNK;
  return c < 1 ? 1 : (c < RW_CHUNK_MAX ? c : RW_CHUNK_MAX);
}
__host__ __device__ inline int conv_rows(int S, int T) { return T + rw_chunks(S) - 1; }
// Whether the launch needs the reduction: kicks of the general block with
// memory (T > 1), or resistive-wall kicks with more than one row.
__host__ __device__ inline bool conv_reduces(int S, int T, bool rw) {
  return T > 1 || (rw && rw_chunks(S) > 1);
}

// Block (tile, ki, t) of nk kick rows: blockIdx.(x, y, z) and gridDim.y of
// the standalone launch, explicit in a batch.
__device__ __forceinline__ void wake_conv_block(const phx_wake_hist *h, const WakeKicks &table,
                                                int S, double *part, int T, double *out,
                                                int gmax, int tile, int ki, int t, int nk) {
  // LDS, convolution form: [WK_TILE + S - 1] wake values, then [S] source
  // moments; dot-product form: [S] source times, then [gmax][S] moments.
  extern __shared__ double vbuf[];
  __shared__ double red[WK_GMAX][WK_THREADS];
  __shared__ double rwp[WK_GMAX][3];
  __shared__ WakeKicks kk;
  stage_kicks(table, nk, kk);
  const phx_wake_kick &w = kk.k[ki];
  if (is_rw_dot(w)) return;  // wake_rw_block's
  const int nf = h->n_filled, head = h->head, cap = h->capacity;
  const int n_turns = nf < w.n_turns_wake ? nf : w.n_turns_wake;
  if (t >= n_turns) return;
  int mem[WK_GMAX];
  int nm = 1;
  mem[0] = ki;
  if (w.dot_product) {
    int pos = 0;
    for (int k = 0; k < ki; ++k) pos += same_shape(kk.k[k], w) ? 1 : 0;
    if (pos % gmax != 0) return;
    for (int k = ki + 1; k < nk && nm < gmax; ++k)
      if (same_shape(kk.k[k], w)) mem[nm++] = k;
  }
  const int j0 = tile * WK_TILE;
  const int jj = threadIdx.x % WK_TILE, g = threadIdx.x / WK_TILE;
  const int j = j0 + jj;
  const int slot = ring_slot(head, t, cap);
  const double age = h->age[slot];
  const double *tsrc = h->times + (size_t)slot * S;
  const double *ttgt = h->times + (size_t)head * S;
  double acc[WK_GMAX];
#pragma unroll
  for (int m = 0; m < WK_GMAX; ++m) acc[m] = 0.0;
  if (w.dot_product) {
    // dt[i, j] = t_target[j] - (t_source[i] + age)     (wake_kicks.py:81-85)
    double *tb = vbuf, *ab = vbuf + S;
    for (int i = threadIdx.x; i < S; i += blockDim.x) tb[i] = tsrc[i] + age;
    for (int m = 0; m < nm; ++m) {
      const double *a = h->mom + ((size_t)slot * 3 + kk.k[mem[m]].moment) * S;
      for (int i = threadIdx.x; i < S; i += blockDim.x) ab[(size_t)m * S + i] = a[i];
      if (threadIdx.x == 0 && w.fn == PHX_WAKE_RESISTIVE_WALL)
        rw_prefixes(kk.k[mem[m]], h->beta[slot], rwp[m]);
    }
    lds_sync();
    if (j < S) {
      const double tj = ttgt[j];
      for (int i = g; i < S; i += WK_GROUPS) {
        const WakeShape b = wake_shape(w, tj - tb[i]);
#pragma unroll
        for (int m = 0; m < WK_GMAX; ++m)
          if (m < nm)
            acc[m] = acc[m] + ab[(size_t)m * S + i] * wake_value(kk.k[mem[m]], b, rwp[m]);
      }
    }
  } else {
    // dt = concat(t_tgt - s_src[-1], (t_tgt - s_src[0])[1:]) and
    // np.convolve(moments, W, 'valid')[j] = sum_m a[m] W[j + S - 1 - m]
    // (wake_kicks.py:109-116)
    double *abuf = vbuf + (WK_TILE + S - 1);
    const double *a = h->mom + ((size_t)slot * 3 + w.moment) * S;
    for (int i = threadIdx.x; i < S; i += blockDim.x) abuf[i] = a[i];
    if (threadIdx.x == 0 && w.fn == PHX_WAKE_RESISTIVE_WALL) rw_prefixes(w, h->beta[slot], rwp[0]);
    lds_sync();
    const double s_last = tsrc[S - 1] + age, s_first = tsrc[0] + age;
    const int p0 = j0, p1 = min(j0 + WK_TILE - 1 + S - 1, 2 * S - 2);
    for (int p = p0 + threadIdx.x; p <= p1; p += blockDim.x) {
      const double dt = p <= S - 1 ? ttgt[p] - s_last : ttgt[p - S + 1] - s_first;
      vbuf[p - p0] = wake_eval(w, dt, rwp[0]);
    }
    lds_sync();
    if (j < S) {
      for (int m = g; m < S; m += WK_GROUPS) acc[0] = acc[0] + abuf[m] * vbuf[j + S - 1 - m - p0];
    }
  }
#pragma unroll
  for (int m = 0; m < WK_GMAX; ++m) red[m][threadIdx.x] = acc[m];
  lds_sync();
  if (g == 0 && j < S) {
    for (int m = 0; m < nm; ++m) {
      double r = red[m][jj];
      for (int q = 1; q < WK_GROUPS; ++q) r = r + red[m][q * WK_TILE + jj];
      const int km = mem[m];
      if (T == 1) out[(size_t)km * S + j] = kk.k[km].wake_factor * (0.0 + r);
      else part[((size_t)km * conv_rows(S, T) + t) * S + j] = r;
    }
  }
}

// Sum of a kick's rows of partial sums, times the wake factor: RED_LANES
// consecutive lanes per (kick, target) each sum a contiguous part of the rows
// in order, and the first adds the parts in order.  Threads:
// RED_LANES * n_kicks * S (blocks of a multiple of 64 threads).
#define RED_LANES 4
__device__ __forceinline__ void wake_conv_reduce(const phx_wake_hist *h, const WakeKicks &kk,
                                                 int n_kicks, int S, const double *part, int T,
                                                 double *out, int bid) {
  const int gidx = bid * blockDim.x + threadIdx.x;
  const int idx = gidx / RED_LANES, sub = gidx % RED_LANES;
  double acc = 0.0, wf = 0.0;
  bool write = false;
  int ki = 0, j = 0;
  if (idx < n_kicks * S) {  // no return before the lane exchange below
    ki = idx / S;
    j = idx % S;
    const phx_wake_kick &w = kk.k[ki];
    if (!(T == 1 && !is_rw_dot(w))) {  // otherwise written by its block
      const int nf = h->n_filled;
      const int n_turns = nf < w.n_turns_wake ? nf : w.n_turns_wake;
      const int rows = is_rw_dot(w) && n_turns > 0 ? n_turns + rw_chunks(S) - 1 : n_turns;
      const int len = (rows + RED_LANES - 1) / RED_LANES;
      const int r0 = min(sub * len, rows), r1 = min(r0 + len, rows);
      const double *p = part + (size_t)ki * conv_rows(S, T) * S + j;
      int q = r0;
      for (; q + 8 <= r1; q += 8) {
        double v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = p[(size_t)(q + r) * S];
#pragma unroll
        for (int r = 0; r < 8; ++r) acc = acc + v[r];
      }
      for (; q < r1; ++q) acc = acc + p[(size_t)q * S];
      write = true;
      wf = w.wake_factor;
    }
  }
  const int lane = threadIdx.x & 63;
  double tot = acc;
#pragma unroll
  for (int r = 1; r < RED_LANES; ++r) tot = tot + __shfl(acc, lane + r);
  if (write && sub == 0) out[(size_t)ki * S + j] = wf * tot;
}

// ---------------------------------------------------------------------------
// Resistive wall in the dot-product form (multi-turn memory).  For dt < -|dt_min|
// the function is W = pre0 * sqrt(k_sigma / dt) * sqrt(Z0 c), pre0 the
// prefix of class 0 (rw_prefixes), k_sigma = -mu_r/(pi sigma) < 0, so
//   W = coef * rsqrt(-dt),    coef = pre0 * sqrt(-k_sigma) * sqrt(Z0 c),
// at dt = -|dt_min| it is half of that and above it zero (pre1 = pre0 / 2
// exactly, pre2 = 0).  The kicks of one shape (same_shape) differ only in
// coef and in the source moment (n, n x, n y), so a block sums
//   acc_u[j] = sum_i m_u[i] * sel(dt_ij) * rsqrt(x_ij)
// once per distinct moment u and writes coef_k * acc_u(k) for every kick k.
// One block = (RW_TILE targets, row of sources: a chunk of the current turn
// or one older turn, conv_rows): the block stages the row's source times and
// moments in LDS once for every shape group, each wave runs over its share
// of the sources (the same source in all lanes: LDS broadcasts) for two
// targets per lane.  Blocks whose every pair has dt < -|dt_min| (all sources
// of turns gone by, decided from the extreme times: rounding is monotonic)
// skip the class selection.  The sums run in a fixed order: each wave in
// source order, then the waves in order.
#define RW_WAVES 8  // the most waves per block (the matrix-core kernel's)
// Waves per block of each kernel kind (KM, see wake_rw_group): the VALU
// kernel runs 4-wave blocks, so that three fit a compute unit (LDS) and every
// block of a 1000-slice, 64-turn convolution is resident at once; the
// matrix-core kernel's tile mapping takes 8 waves of 16 targets.
template <int KM>
__host__ __device__ constexpr int rw_waves() {
  return KM == 0 ? 4 : 8;
}
template <int KM>
__host__ __device__ constexpr int rw_threads() {
  return rw_waves<KM>() * 64;
}
#define RW_JT 2
#define RW_TILE (RW_JT * 64)

// 1/sqrt(x), x > 0: the hardware estimate y (relative error e0 < 2^-22) and
// one third-order step y (1 + e/2 + 3e^2/8), e = 1 - x y^2, which leaves
// O(e0^3) plus rounding: within a few ulp.
__device__ __forceinline__ double rw_rsqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x, y * y, 1.0);
  return fma(y, e * fma(0.375, e, 0.5), y);
}

// The class-0 prefix, rw_prefixes' chain for c = 0.
__device__ __forceinline__ double rw_prefix0(const phx_wake_kick &w, double beta_src) {
  double t = w.yokoya * -2.0;
  t = t / 2.0;
  t = t * sqrt(beta_src);
  t = t * w.L;
  t = t / PHX_PI;
  return t / w.b3;
}

template <bool UNIFORM>
__device__ __forceinline__ double rw_pair(double tja, double ts, double dmin) {
  const double dt = tja - ts;
  if (UNIFORM) return rw_rsqrt(-dt);
  // dt.clip(max=-|dt_min|) and the class of sign(dt + |dt_min|)
  const double a = dt + dmin;
  const double x = dt > -dmin ? dmin : -dt;
  const double sel = a < 0.0 ? 1.0 : (a == 0.0 ? 0.5 : (a > 0.0 ? 0.0 : a));
  return sel * rw_rsqrt(x);
}

// Sources [ib, ie) of the LDS arrays, in order, for RW_JT targets.
template <int NU, bool UNIFORM, bool F32>
__device__ __forceinline__ void rw_sum(const double *ts, const double *m0, const double *m1,
                                       const double *m2, int ib, int ie,
                                       const double (&tja)[RW_JT], double dmin,
                                       double (&acc)[3][RW_JT]) {
  if (!F32) {
#pragma unroll 4
    for (int i = ib; i < ie; ++i) {
      const double t = ts[i];
      const double a0 = m0[i], a1 = NU > 1 ? m1[i] : 0.0, a2 = NU > 2 ? m2[i] : 0.0;
#pragma unroll
      for (int q = 0; q < RW_JT; ++q) {
        const double w = rw_pair<UNIFORM>(tja[q], t, dmin);
        acc[0][q] = fma(a0, w, acc[0][q]);
        if (NU > 1) acc[1][q] = fma(a1, w, acc[1][q]);
        if (NU > 2) acc[2][q] = fma(a2, w, acc[2][q]);
      }
    }
    return;
  }
  // FP32: the wake value rsqrt(x) from x rounded to FP32 (the difference dt
  // itself is taken in FP64), the moments rounded to FP32, FP32 sums
  float fa[3][RW_JT] = {};
#pragma unroll 4
  for (int i = ib; i < ie; ++i) {
    const double t = ts[i];
    const float a0 = (float)m0[i], a1 = NU > 1 ? (float)m1[i] : 0.f, a2 = NU > 2 ? (float)m2[i] : 0.f;
#pragma unroll
    for (int q = 0; q < RW_JT; ++q) {
      const double dt = tja[q] - t;
      float w;
      if (UNIFORM) {
        w = __builtin_amdgcn_rsqf((float)(-dt));
      } else {
        const double a = dt + dmin;
        const double x = dt > -dmin ? dmin : -dt;
        const float sel = a < 0.0 ? 1.f : (a == 0.0 ? 0.5f : (a > 0.0 ? 0.f : (float)a));
        w = sel * __builtin_amdgcn_rsqf((float)x);
      }
      fa[0][q] = fmaf(a0, w, fa[0][q]);
      if (NU > 1) fa[1][q] = fmaf(a1, w, fa[1][q]);
      if (NU > 2) fa[2][q] = fmaf(a2, w, fa[2][q]);
    }
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int q = 0; q < RW_JT; ++q) acc[u][q] = (double)fa[u][q];
}

template <bool F32>
__device__ __forceinline__ void rw_dispatch(int nu, bool uniform, const double *ts,
                                            const double *m0, const double *m1,
                                            const double *m2, int ib, int ie,
                                            const double (&tja)[RW_JT], double dmin,
                                            double (&acc)[3][RW_JT]) {
  if (uniform) {
    if (nu == 1) rw_sum<1, true, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
    else if (nu == 2) rw_sum<2, true, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
    else rw_sum<3, true, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
  } else {
    if (nu == 1) rw_sum<1, false, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
    else if (nu == 2) rw_sum<2, false, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
    else rw_sum<3, false, F32>(ts, m0, m1, m2, ib, ie, tja, dmin, acc);
  }
}

// ---------------------------------------------------------------------------
// Source turns gone by: separable expansion.  For a past turn every pair has
// tau = -dt = ts_i - tja_j > |dt_min| and tau = A + u_i - v_j with
// A = c_s - c_t the distance of the two centres (about one revolution
// period per turn of age) and |u_i| + |v_j| <= r A, r <= RW_EXP_R.  Then
//   tau^(-1/2) = A^(-1/2) sum_k c_k ((u_i - v_j)/A)^k,  c_k = binom(-1/2, k),
// and sum_i m_i tau_ij^(-1/2) = A^(-1/2) sum_q G_q (-y_j)^q with
// y_j = v_j / A, G_q = sum_p c_(p+q) C(p+q, p) U_p, U_p = sum_i m_i (u_i/A)^p:
// O(S) per turn instead of O(S^2).  With RW_EXP_K = 6 terms the truncation
// is below c_6 r^6 / (1 - r) < 2.3e-19 of sum_i |m_i| tau_ij^(-1/2); the sums
// themselves round like the pair sums (about 1e-16 relative).
// ---------------------------------------------------------------------------
#define RW_EXP_K 6
#define RW_EXP_R 1e-3

__device__ __forceinline__ double rw_binom_c(int k) {  // binom(-1/2, k), exact
  constexpr double c[RW_EXP_K] = {1.0, -0.5, 0.375, -0.3125, 0.2734375, -0.24609375};
  return c[k];
}

__device__ __forceinline__ double rw_choose(int n, int k) {
  constexpr double t[RW_EXP_K][RW_EXP_K] = {{1, 0, 0, 0, 0, 0},   {1, 1, 0, 0, 0, 0},
                                           {1, 2, 1, 0, 0, 0},   {1, 3, 3, 1, 0, 0},
                                           {1, 4, 6, 4, 1, 0},   {1, 5, 10, 10, 5, 1}};
  return t[n][k];
}

// ---------------------------------------------------------------------------
// Matrix-core pair loops (PHX_PREC_MFMA_F32, PHX_PREC_MFMA_BF16X3).  One
// 4x4x1 (f32) or 4x4x4 (bf16) MFMA of 16 blocks covers 16 sources x 4
// targets: lane l takes source s = i0 + l/4 of block l/4 and target
// j = 4q + l%4 of its quad; the A operand is the moment of row l%4 of source
// s (rows: the group's distinct moments, zero beyond), the B operand the wake
// value W(s, j), so block b accumulates D[moment][target] over its sources.
// Each wave owns 4 quads (16 targets) and runs over every source; the FP32
// accumulators are added into FP64 every RW_MF_FLUSH steps (256 sources),
// and the 16 blocks are summed in FP64 by lane exchanges in a fixed order.
// bf16x3: W = W1 + W2 and m = m1 + m2 in bfloat16 (round to nearest even,
// each part the rounded remainder of the previous ones); one 4-deep product
// m1 W1 + m2 W1 + m1 W2 (about 16 significant bits).  bf16x6: three parts
// each and two 4-deep products, m1 W1 + m1 W2 + m2 W1 + m1 W3 and
// m2 W2 + m3 W1 (every term down to 2^-16 of the leading one).
// ---------------------------------------------------------------------------
typedef float rw_f4 __attribute__((ext_vector_type(4)));
typedef short rw_s4 __attribute__((ext_vector_type(4)));
#define RW_MF_QUADS 4
#define RW_MF_FLUSH 16

__device__ __forceinline__ short rw_bf16_bits(float v) {
  return __builtin_bit_cast(short, (__bf16)v);
}

__device__ __forceinline__ float rw_bf16_value(short b) {
  return __builtin_bit_cast(float, (unsigned)(unsigned short)b << 16);
}

// MF: 0 FP32, 1 bf16x3, 2 bf16x6
template <bool UNIFORM, int MF>
__device__ __forceinline__ void rw_mfma(const double *ts, const double *m0, const double *m1,
                                        const double *m2, int S, int nu,
                                        const double (&tq)[RW_MF_QUADS], double dmin,
                                        double (&accd)[RW_MF_QUADS][4]) {
  const int lane = threadIdx.x & 63;
  const int row = lane & 3, sub = lane >> 2;
  const double *mrow = row == 0 ? m0 : (row == 1 ? m1 : m2);
  const bool live_row = row < nu;
  rw_f4 acc[RW_MF_QUADS];
#pragma unroll
  for (int q = 0; q < RW_MF_QUADS; ++q) acc[q] = rw_f4{0.f, 0.f, 0.f, 0.f};
  int steps = 0;
  for (int i0 = 0; i0 < S; i0 += 16) {
    const int src = i0 + sub;
    const bool valid = src < S;
    const double t = valid ? ts[src] : 0.0;
    const float a = (valid && live_row) ? (float)mrow[src] : 0.f;
    rw_s4 av, av2;
    if (MF > 0) {
      const short a1 = rw_bf16_bits(a);
      const float ar = a - rw_bf16_value(a1);
      const short a2 = rw_bf16_bits(ar);
      if (MF == 1) {
        av = rw_s4{a1, a2, a1, 0};
      } else {
        const short a3 = rw_bf16_bits(ar - rw_bf16_value(a2));
        av = rw_s4{a1, a1, a2, a1};
        av2 = rw_s4{a2, a3, 0, 0};
      }
    }
#pragma unroll
    for (int q = 0; q < RW_MF_QUADS; ++q) {
      const double dt = tq[q] - t;
      float w;
      if (UNIFORM) {
        w = __builtin_amdgcn_rsqf((float)(-dt));
      } else {
        const double g = dt + dmin;
        const double x = dt > -dmin ? dmin : -dt;
        const float sel = g < 0.0 ? 1.f : (g == 0.0 ? 0.5f : (g > 0.0 ? 0.f : (float)g));
        w = sel * __builtin_amdgcn_rsqf((float)x);
      }
      w = valid ? w : 0.f;
      if (MF > 0) {
        const short w1 = rw_bf16_bits(w);
        const float wr = w - rw_bf16_value(w1);
        const short w2 = rw_bf16_bits(wr);
        if (MF == 1) {
          const rw_s4 bv = rw_s4{w1, w1, w2, 0};
          acc[q] = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(av, bv, acc[q], 0, 0, 0);
        } else {
          const short w3 = rw_bf16_bits(wr - rw_bf16_value(w2));
          const rw_s4 bv = rw_s4{w1, w2, w1, w3};
          const rw_s4 bv2 = rw_s4{w2, w1, 0, 0};
          acc[q] = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(av, bv, acc[q], 0, 0, 0);
          acc[q] = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(av2, bv2, acc[q], 0, 0, 0);
        }
      } else {
        acc[q] = __builtin_amdgcn_mfma_f32_4x4x1f32(a, w, acc[q], 0, 0, 0);
      }
    }
    if (++steps == RW_MF_FLUSH) {
      steps = 0;
#pragma unroll
      for (int q = 0; q < RW_MF_QUADS; ++q) {
#pragma unroll
        for (int r = 0; r < 4; ++r) accd[q][r] = accd[q][r] + (double)acc[q][r];
        acc[q] = rw_f4{0.f, 0.f, 0.f, 0.f};
      }
    }
  }
#pragma unroll
  for (int q = 0; q < RW_MF_QUADS; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) accd[q][r] = accd[q][r] + (double)acc[q][r];
  // sum of the 16 blocks: lanes of one target column differ in bits 2..5
#pragma unroll
  for (int off = 4; off < 64; off <<= 1)
#pragma unroll
    for (int q = 0; q < RW_MF_QUADS; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) accd[q][r] = accd[q][r] + __shfl_xor(accd[q][r], off);
}

__device__ __forceinline__ double wave_min_nan(double v, int &bad) {
  bad |= v != v;
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// Whether the precision runs on the matrix cores.
__host__ __device__ __forceinline__ bool rw_is_mfma(int precision) {
  return precision == PHX_PREC_MFMA_F32 || precision == PHX_PREC_MFMA_BF16X3 ||
         precision == PHX_PREC_MFMA_BF16X6;
}

// Description of a launch's kicks for the resistive-wall blocks, built once
// per convolution (rw_desc_build) ahead of them: per kick its group (the
// first kick of its shape), the moments the group reads, the rank of its
// moment among them, its flags, memory depth, precision, |dt_min| and wake
// factor; then per source turn t < T (age index) and kick k the coefficient
// pre0 sqrt(-k_sigma) sqrt(Z0 c) at that turn's beta, at coef[t * nk + k].
// The blocks read it with one load each, instead of walking the kick table
// for the groups and evaluating the coefficients themselves.
#define RWD_RW 1    // resistive wall in the dot-product form
#define RWD_MFMA 2  // on the matrix cores
struct RwKick {
  int lead, mask, slot, flags, n_turns, precision;
  double dmin, wake_factor;
};
static_assert(sizeof(RwKick) % sizeof(uint64_t) == 0, "kick descriptions are 8-byte words");
__host__ __device__ inline size_t rw_desc_bytes(int T) {
  return PHX_MAX_KICKS * sizeof(RwKick) + (size_t)(T > 0 ? T : 1) * PHX_MAX_KICKS * sizeof(double);
}
__host__ __device__ inline double *rw_desc_coef(RwKick *d) {
  return reinterpret_cast<double *>(d + PHX_MAX_KICKS);
}
__host__ __device__ inline const double *rw_desc_coef(const RwKick *d) {
  return reinterpret_cast<const double *>(d + PHX_MAX_KICKS);
}

// Builds the description of the nk kicks `table` over T turns of history h;
// called by every thread of one block (any size).
__device__ __forceinline__ void rw_desc_build(const phx_wake_hist *h, const WakeKicks &table,
                                              int nk, int T, RwKick *d) {
  __shared__ WakeKicks kk;
  stage_kicks(table, nk, kk);
  if (threadIdx.x < nk) {
    const int k = threadIdx.x;
    const phx_wake_kick &kw = kk.k[k];
    int lead = k;
    unsigned mask = 0;
    for (int m = nk - 1; m >= 0; --m) {
      if (!same_shape(kk.k[m], kw)) continue;
      mask |= 1u << kk.k[m].moment;
      lead = m < lead ? m : lead;
    }
    RwKick r;
    r.lead = lead;
    r.mask = (int)mask;
    r.slot = __popc(mask & ((1u << kw.moment) - 1u));
    r.flags = (is_rw_dot(kw) ? RWD_RW : 0) | (rw_is_mfma(kw.precision) ? RWD_MFMA : 0);
    r.n_turns = kw.n_turns_wake;
    r.precision = kw.precision;
    r.dmin = kw.dt_min_abs;
    r.wake_factor = kw.wake_factor;
    d[k] = r;
  }
  const int nf = h->n_filled, head = h->head, cap = h->capacity;
  double *coef = rw_desc_coef(d);
  for (int i = threadIdx.x; i < T * nk; i += blockDim.x) {
    const int t = i / nk, k = i % nk;
    const phx_wake_kick &kw = kk.k[k];
    double c = 0.0;
    if (t < nf && is_rw_dot(kw)) {
      const double beta = h->beta[ring_slot(head, t, cap)];
      c = (rw_prefix0(kw, beta) * sqrt(-kw.k_sigma)) * kw.sqrt_z0c;
    }
    coef[i] = c;
  }
}

// Static LDS of wake_rw_block: the kick descriptions, the block's turn's
// coefficients and the waves' extremes.
#define RW_STATIC_LDS                                                     \
  (PHX_MAX_KICKS * (sizeof(RwKick) + sizeof(double)) +                    \
   RW_WAVES * (4 * sizeof(double) + sizeof(int)))
// Parts of the source range in the power sums of the past-turn expansion
// (one thread each).
#define RW_UP_PARTS 128
#define RW_UP_SUMS (3 * RW_EXP_K)
#define RW_UP_LD (RW_UP_PARTS + 1)  // row stride of the parts (LDS banks)
// Reduction area: the waves' sums of one moment (pair loops), or the parts of
// the power sums and then the sums themselves (expansion).
#define RW_RED_DOUBLES                                                        \
  (RW_WAVES * RW_TILE > RW_UP_SUMS * (RW_UP_LD + 1) ? RW_WAVES * RW_TILE      \
                                                     : RW_UP_SUMS * (RW_UP_LD + 1))
// Dynamic LDS of wake_rw_block: the source turn's times and its three
// moments (staged once for every group of the block), then the reduction
// area.
inline size_t wake_rw_lds_bytes(int S) {
  return ((size_t)4 * S + (size_t)RW_RED_DOUBLES) * sizeof(double);
}

// What wake_rw_block stages for its groups: the sources [sb, sb + n) of
// source turn t (ring slot `slot`) at the start of the dynamic LDS (times,
// then moments n, n x, n y, S apart), the lane's RW_JT targets as
// t_target - age, and the extremes of the sources and of the tile's targets
// over the block, with a NaN flag.
struct RwStage {
  int t, row, n, head;
  double age;
  double tja[RW_JT];
  double mn, smax, tgmin, tgmax;
  int bad;
};

// The group of kick ki (its first kick) for targets [tile * RW_TILE, +RW_TILE)
// and the staged sources, into row st.row of R; called by all threads of the
// block.  dk: the kick descriptions; coef[k]: kick k's coefficient at the
// staged turn.
// KM: the kernel's kind, 0 for the VALU precisions (FP64 with the past-turn
// expansion, FP64 pair by pair, FP32), 1 for the matrix-core ones; each
// kind is its own kernel, so that neither carries the other's registers.

template <int KM>
__device__ __forceinline__ void wake_rw_group(const phx_wake_hist *h, const RwKick *dk, int nk,
                                              int S, double *part, int R, double *out, int ki,
                                              int tile, const RwStage &st, const double *coef) {
  extern __shared__ double rw_lds[];
  const RwKick &w = dk[ki];
  const unsigned mask = (unsigned)w.mask;
  const int nu = __popc(mask);
  const int mo0 = __ffs(mask) - 1;
  const int mo1 = nu > 1 ? __ffs(mask & ~(1u << mo0)) - 1 : mo0;
  const int mo2 = nu > 2 ? 2 : mo1;

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int n = st.n;
  const double *ts = rw_lds;
  const double *m0 = rw_lds + (size_t)(1 + mo0) * S, *m1 = rw_lds + (size_t)(1 + mo1) * S,
               *m2 = rw_lds + (size_t)(1 + mo2) * S;
  double *red = rw_lds + (size_t)4 * S;
  // the extremes of the tile's targets: every wave holds all of them
  const bool uniform = !st.bad && (st.tgmax - st.mn) + w.dmin < 0.0;
  const double dmin = w.dmin;
  double tot[3] = {0.0, 0.0, 0.0};
  // the past-turn expansion (FP64 precision): centres, their distance A and
  // the largest relative offset r = (max|u| + max|v|) / A
  const double c_s = 0.5 * (st.mn + st.smax), c_t = 0.5 * (st.tgmin + st.tgmax);
  const double A = c_s - c_t;
  const double r_off = ((st.smax - c_s) + (st.tgmax - c_t)) / A;
  const bool expand = w.precision == PHX_PREC_FP64 && uniform && st.t > 0 && A > 0.0 &&
                      r_off <= RW_EXP_R;
  if (KM == 0 && expand) {
    // U_p of each distinct moment: thread r < P sums part r of the sources in
    // source order, every power of every moment (term m x^p as ((m x) x)...);
    // thread q < 3 K then adds the parts of sum q in order
    const double inv_A = 1.0 / A;
    if (threadIdx.x < RW_UP_PARTS) {
      const int len = (n + RW_UP_PARTS - 1) / RW_UP_PARTS;
      const int ib = min((int)threadIdx.x * len, n), ie = min(ib + len, n);
      double U[3][RW_EXP_K] = {};
      for (int i = ib; i < ie; ++i) {
        const double x = (ts[i] - c_s) * inv_A;
        const double mm[3] = {m0[i], nu > 1 ? m1[i] : 0.0, nu > 2 ? m2[i] : 0.0};
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          double pw = mm[u];
#pragma unroll
          for (int p = 0; p < RW_EXP_K; ++p) {
            U[u][p] = U[u][p] + pw;
            pw = pw * x;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 3; ++u)
#pragma unroll
        for (int p = 0; p < RW_EXP_K; ++p)
          red[(u * RW_EXP_K + p) * RW_UP_LD + threadIdx.x] = U[u][p];
    }
    lds_sync();
    double *usum = red + RW_UP_SUMS * RW_UP_LD;
    if (threadIdx.x < RW_UP_SUMS) {
      const double *pr = red + threadIdx.x * RW_UP_LD;
      double v = pr[0];
#pragma unroll 8
      for (int q = 1; q < RW_UP_PARTS; ++q) v = v + pr[q];
      usum[threadIdx.x] = v;
    }
    lds_sync();
    if (threadIdx.x < RW_TILE) {
      const double tj = wave == 0 ? st.tja[0] : st.tja[RW_JT - 1];
      const double ny = -((tj - c_t) * inv_A);
      const double scale = 1.0 / sqrt(A);
      for (int u = 0; u < nu; ++u) {
        double Up[RW_EXP_K];
#pragma unroll
        for (int p = 0; p < RW_EXP_K; ++p) Up[p] = usum[u * RW_EXP_K + p];
        double f = 0.0;  // Horner in -y over q = K-1 .. 0
#pragma unroll
        for (int q = RW_EXP_K - 1; q >= 0; --q) {
          double g = 0.0;
#pragma unroll
          for (int p = 0; p + q < RW_EXP_K; ++p)
            g = g + (rw_binom_c(p + q) * rw_choose(p + q, p)) * Up[p];
          f = f * ny + g;
        }
        tot[u] = scale * f;
      }
    }
  } else if (KM == 1) {
    // wave w: targets [16 w, 16 w + 16) of the tile, in quads
    double tq[RW_MF_QUADS];
#pragma unroll
    for (int q = 0; q < RW_MF_QUADS; ++q) {
      const int jt = tile * RW_TILE + wave * 4 * RW_MF_QUADS + 4 * q + (lane & 3);
      tq[q] = h->times[(size_t)st.head * S + (jt < S ? jt : S - 1)] - st.age;
    }
    double accd[RW_MF_QUADS][4] = {};
    const int mf = w.precision == PHX_PREC_MFMA_F32 ? 0 : (w.precision == PHX_PREC_MFMA_BF16X3 ? 1 : 2);
    if (uniform) {
      if (mf == 0) rw_mfma<true, 0>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
      else if (mf == 1) rw_mfma<true, 1>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
      else rw_mfma<true, 2>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
    } else {
      if (mf == 0) rw_mfma<false, 0>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
      else if (mf == 1) rw_mfma<false, 1>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
      else rw_mfma<false, 2>(ts, m0, m1, m2, n, nu, tq, dmin, accd);
    }
    if (lane < 4) {
#pragma unroll
      for (int q = 0; q < RW_MF_QUADS; ++q)
#pragma unroll
        for (int u = 0; u < 3; ++u)
          red[u * RW_TILE + wave * 4 * RW_MF_QUADS + 4 * q + lane] = accd[q][u];
    }
    lds_sync();
    if (threadIdx.x < RW_TILE)
      for (int u = 0; u < nu; ++u) tot[u] = red[u * RW_TILE + threadIdx.x];
  } else {
    constexpr int WV = rw_waves<KM>();
    const int chunk = (n + WV - 1) / WV;
    const int ib = min(wave * chunk, n), ie = min(ib + chunk, n);
    double acc[3][RW_JT] = {};
    if (w.precision == PHX_PREC_FP32)
      rw_dispatch<true>(nu, uniform, ts, m0, m1, m2, ib, ie, st.tja, dmin, acc);
    else
      rw_dispatch<false>(nu, uniform, ts, m0, m1, m2, ib, ie, st.tja, dmin, acc);
    // the reduction area takes the wave sums, one moment at a time; thread
    // r < RW_TILE finishes target j0 + r
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      if (u >= nu) break;
      if (u > 0) lds_sync();
#pragma unroll
      for (int q = 0; q < RW_JT; ++q) red[wave * RW_TILE + q * 64 + lane] = acc[u][q];
      lds_sync();
      if (threadIdx.x < RW_TILE) {
        double r = red[threadIdx.x];
        for (int q = 1; q < WV; ++q) r = r + red[q * RW_TILE + threadIdx.x];
        tot[u] = r;
      }
    }
  }
  const int j = tile * RW_TILE + threadIdx.x;
  if (threadIdx.x < RW_TILE && j < S) {
    for (int k = ki; k < nk; ++k) {
      if (dk[k].lead != ki) continue;
      const int u = dk[k].slot;
      const double sum = u == 0 ? tot[0] : (u == 1 ? tot[1] : tot[2]);
      const double v = coef[k] * sum;
      if (R == 1) out[(size_t)k * S + j] = dk[k].wake_factor * (0.0 + v);
      else part[((size_t)k * R + st.row) * S + j] = v;
    }
  }
  lds_sync();  // the reduction area is the next group's
}

// One block = (RW_TILE targets, row): blockIdx.x = tile + tiles * row
// (conv_rows(S, T) rows: the chunks of turn 0, then turns 1..T-1), so that
// consecutive blocks (which go to different XCDs) cover different rows.  The
// block stages the row's sources once and serves every resistive-wall group
// (in kick order) whose memory reaches the row's turn.  The ring's loads
// (its state, then the row's sources, the targets and the slot's age) are
// issued before the kick description is read, each dependent load being a
// round trip to memory that the next one waits for.  desc: the kicks'
// description (rw_desc_build).
template <int KM>
__device__ __forceinline__ void wake_rw_block(const phx_wake_hist *h, const RwKick *desc,
                                              int nk, int S, double *part, int T, double *out,
                                              int bid) {
  extern __shared__ double rw_lds[];
  __shared__ RwKick dk[PHX_MAX_KICKS];
  __shared__ double coef[PHX_MAX_KICKS];
  __shared__ double wmin[RW_WAVES], wmax[RW_WAVES], wtmax[RW_WAVES], wtmin[RW_WAVES];
  __shared__ int wbad[RW_WAVES];
  const int head = h->head, cap = h->capacity, nf = h->n_filled;
  const int tiles = (S + RW_TILE - 1) / RW_TILE;
  const int nc = rw_chunks(S), R = T + nc - 1;
  const int tile = bid % tiles, row = bid / tiles;
  RwStage st;
  st.t = row < nc ? 0 : row - nc + 1;
  st.row = row;
  st.head = head;
  if (st.t >= nf) return;  // beyond every group's memory (the whole block)
  const int cs = (S + nc - 1) / nc;
  const int sb = row < nc ? min(row * cs, S) : 0;
  st.n = (row < nc ? min(sb + cs, S) : S) - sb;
  const int slot = ring_slot(head, st.t, cap);
  const double age = h->age[slot];
  st.age = age;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const double *tsrc = h->times + (size_t)slot * S + sb;
  const double *mom = h->mom + (size_t)slot * 3 * S + sb;
  int bad = 0;
  double mn = INFINITY, mxs = -INFINITY;
  for (int i = threadIdx.x; i < st.n; i += rw_threads<KM>()) {
    const double v = tsrc[i];
    rw_lds[i] = v;
    bad |= v != v;
    mn = v < mn ? v : mn;
    mxs = v > mxs ? v : mxs;
    rw_lds[S + i] = mom[i];
    rw_lds[2 * S + i] = mom[(size_t)S + i];
    rw_lds[3 * S + i] = mom[(size_t)2 * S + i];
  }
  // dt = (t_target - age) - t_source: the age leaves the pair loop
  double mx = -INFINITY, mnt = INFINITY;
#pragma unroll
  for (int q = 0; q < RW_JT; ++q) {
    const int j = tile * RW_TILE + q * 64 + lane;
    st.tja[q] = h->times[(size_t)head * S + (j < S ? j : S - 1)] - age;
    bad |= st.tja[q] != st.tja[q];
    mx = st.tja[q] > mx ? st.tja[q] : mx;
    mnt = st.tja[q] < mnt ? st.tja[q] : mnt;
  }
  mn = wave_min_nan(mn, bad);
  mxs = -wave_min_nan(-mxs, bad);
  const double tmax = -wave_min_nan(-mx, bad);
  const double tmin = wave_min_nan(mnt, bad);
  const int wave_bad = __ballot(bad) != 0ULL;
  if (lane == 0) {
    wmin[wave] = mn;
    wmax[wave] = mxs;
    wtmax[wave] = tmax;
    wtmin[wave] = tmin;
    wbad[wave] = wave_bad;
  }
  {  // the kick descriptions and the turn's coefficients
    constexpr int words = (int)(sizeof(RwKick) / sizeof(uint64_t));
    const uint64_t *src = reinterpret_cast<const uint64_t *>(desc);
    uint64_t *dst = reinterpret_cast<uint64_t *>(dk);
    for (int i = threadIdx.x; i < nk * words; i += rw_threads<KM>()) dst[i] = src[i];
    if (threadIdx.x < nk) coef[threadIdx.x] = rw_desc_coef(desc)[st.t * nk + threadIdx.x];
  }
  lds_sync();  // the sources, the waves' extremes and the kick descriptions
  st.bad = 0;
  st.mn = mn;
  st.smax = mxs;
  st.tgmax = tmax;
  st.tgmin = tmin;
  for (int q = 0; q < rw_waves<KM>(); ++q) {
    st.mn = wmin[q] < st.mn ? wmin[q] : st.mn;
    st.smax = wmax[q] > st.smax ? wmax[q] : st.smax;
    st.tgmax = wtmax[q] > st.tgmax ? wtmax[q] : st.tgmax;
    st.tgmin = wtmin[q] < st.tgmin ? wtmin[q] : st.tgmin;
    st.bad |= wbad[q];
  }
  for (int ki = 0; ki < nk; ++ki) {
    const RwKick &w = dk[ki];
    if (w.lead != ki || !(w.flags & RWD_RW) || ((w.flags & RWD_MFMA) != 0) != (KM == 1)) continue;
    const int n_turns = nf < w.n_turns ? nf : w.n_turns;
    if (st.t >= n_turns) continue;
    wake_rw_group<KM>(h, dk, nk, S, part, R, out, ki, tile, st, coef);
  }
}

// Static LDS of wake_conv_block; its dynamic LDS serves up to gmax kicks per
// block in the dot-product form.
#define WK_STATIC_LDS ((WK_GMAX * WK_THREADS + 3 * WK_GMAX) * sizeof(double) + sizeof(WakeKicks))
inline size_t wake_conv_lds_bytes(int S, int gmax) {
  const size_t conv = (size_t)(WK_TILE + 2 * S - 1), dot = (size_t)(1 + gmax) * S;
  return (conv > dot ? conv : dot) * sizeof(double);
}
// The most kicks per block (<= WK_GMAX) whose moments fit the LDS, at least 1.
inline int wake_conv_gmax(int S, size_t max_lds) {
  for (int gm = WK_GMAX; gm > 1; --gm)
    if (wake_conv_lds_bytes(S, gm) + WK_STATIC_LDS <= max_lds) return gm;
  return 1;
}
