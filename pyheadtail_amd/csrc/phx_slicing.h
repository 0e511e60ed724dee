// Slicing building blocks shared by the standalone slicing calls and the
// fused engine: the z-moment block that fixes the cuts, and the per-slice
// fixed-point accumulators.
#pragma once

#include "phx_device.h"

size_t phx_max_lds_bytes();
int phx_num_cus();

struct Coords6 {
  const double *c[6];
};

// ---------------------------------------------------------------------------
// z block: [0] sum z  [1] sum (z - z0)  [2] sum (z - z0)^2  [3] min z
//          [4] max z  [5..10] max |u| over finite values of x,xp,y,yp,z,dp
// ---------------------------------------------------------------------------
#define ZB_NV 11
#define ZB_MINM (1ULL << 3)
#define ZB_MAXM (0x7FULL << 4)

// Bit i of AMASK: keep max |u_i| (the others stay -inf).  Comparisons
// instead of fmin/fmax: the same results here (a NaN z or u never replaces
// the running value, which is never NaN) without the NaN quieting of the
// fmin/fmax instructions.
template <unsigned AMASK = 0x3Fu>
__device__ __forceinline__ void zblock_accumulate(double (&v)[ZB_NV], const P6 &p,
                                                  double shift) {
  const double dz = p.z - shift;
  v[0] = v[0] + p.z;
  v[1] = v[1] + dz;
  v[2] = v[2] + dz * dz;
  v[3] = p.z < v[3] ? p.z : v[3];
  v[4] = p.z > v[4] ? p.z : v[4];
  const double a[6] = {fabs(p.x), fabs(p.xp), fabs(p.y), fabs(p.yp), fabs(p.z), fabs(p.dp)};
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if (((AMASK >> i) & 1u) && a[i] < INFINITY && a[i] > v[5 + i]) v[5 + i] = a[i];
}

// Called by one block with the reduced z block in res (shared): writes the
// cuts into *state and the n_slices+1 bin edges into z_bins.
__device__ __forceinline__ void zblock_finalize(const double *res, const phx_slicer &cfg,
                                                int64_t n, phx_slice_state *state,
                                                double *z_bins, int turn) {
  __shared__ double sh_tail, sh_width;
  if (threadIdx.x == 0) {
    phx_slice_state st;
    compute_cuts(cfg, n, res[0], res[1], res[2], res[3], res[4], st);
    for (int i = 0; i < 6; ++i) st.absmax[i] = res[5 + i];
    st.turn = turn;
    st.pad = 0;
    *state = st;
    sh_tail = st.tail;
    sh_width = st.width;
  }
  lds_sync();
  for (int i = threadIdx.x; i <= cfg.n_slices; i += blockDim.x)
    z_bins[i] = arange_value(sh_tail, sh_width, i);
  lds_sync();
}

// ---------------------------------------------------------------------------
// Per-slice accumulation terms and fixed-point accumulators.
// ---------------------------------------------------------------------------
struct AccTerms {
  int n;
  int a[PHX_MAX_TERMS], b[PHX_MAX_TERMS];
  const double *ca[PHX_MAX_TERMS];
  const double *cb[PHX_MAX_TERMS];
};

// Per-block partial histograms (no cross-block atomics: device-scope atomics
// to one address from many blocks serialise, so each block stores its own
// histogram and a slice-parallel kernel merges them).  Stored in tiles of
// ACC_TW consecutive slices, every tile holding all H partials, so that a
// merge block (one or two tiles) reads only whole cache lines of its own:
//   cnt[tile][h][ACC_TW] (u32), hl[tile][h][t][ACC_TW] = {hi, lo} (u64
//   pairs; one 128-byte line per tile and partial for t = 2),
//   fallback[t][s] (f64, only for non-finite or out-of-scale values, merged
//   with FP64 atomics).
#define ACC_TW 4

__host__ __device__ inline size_t acc_cnt_index(int H, int h, int s) {
  return ((size_t)(s / ACC_TW) * H + h) * ACC_TW + s % ACC_TW;
}
__host__ __device__ inline size_t acc_hl_index(int H, int T, int h, int i, int s) {
  return (((size_t)(s / ACC_TW) * H + h) * T + i) * ACC_TW + s % ACC_TW;
}
__host__ __device__ inline size_t acc_tiled(int S) {  // S rounded up to whole tiles
  return ((size_t)S + ACC_TW - 1) / ACC_TW * ACC_TW;
}

struct AccParts {
  unsigned int *cnt;
  ulonglong2 *hl;
  double *fallback;
  int H;  // number of partial histograms (blocks)
};

struct AccLds {
  unsigned long long *hi;
  unsigned long long *lo;
  unsigned int *counts;
};

// Standalone workspace: [256 B reserved][chunk partials][partial histograms]
inline size_t acc_partials_bytes(int64_t n) {
  return ((size_t)phx_pass_grid(n) * 64 * sizeof(double) + 255) / 256 * 256;
}

__host__ __device__ inline size_t acc_parts_bytes(int H, int S, int T) {
  const size_t St = acc_tiled(S);
  return ((size_t)H * St * 4 + 255) / 256 * 256 + (size_t)H * T * St * 16 + (size_t)T * S * 8;
}

__host__ __device__ inline AccParts acc_parts_layout(void *base, int H, int S, int T) {
  const size_t St = acc_tiled(S);
  char *p = reinterpret_cast<char *>(base);
  AccParts a;
  a.H = H;
  a.cnt = reinterpret_cast<unsigned int *>(p);
  p += ((size_t)H * St * 4 + 255) / 256 * 256;
  a.hl = reinterpret_cast<ulonglong2 *>(p);
  p += (size_t)H * T * St * 16;
  a.fallback = reinterpret_cast<double *>(p);
  return a;
}

inline size_t acc_lds_bytes(int S, int T) { return (size_t)T * S * 16 + (size_t)S * 4; }

__device__ __forceinline__ AccLds acc_lds_layout(unsigned long long *base, int S, int T) {
  AccLds L;
  L.hi = base;
  L.lo = base + (size_t)T * S;
  L.counts = reinterpret_cast<unsigned int *>(base + (size_t)2 * T * S);
  return L;
}

__device__ __forceinline__ void acc_lds_zero(const AccLds &L, int S, int T) {
  for (int i = threadIdx.x; i < T * S; i += blockDim.x) {
    L.hi[i] = 0ULL;
    L.lo[i] = 0ULL;
  }
  for (int i = threadIdx.x; i < S; i += blockDim.x) L.counts[i] = 0u;
  lds_sync();
}

// Bound of |term| from max|u| (centres are per-slice means, |c| <= max|u|).
__device__ __forceinline__ int term_exponent(const AccTerms &t, int i, const double *absmax) {
  double ba = absmax[t.a[i]] * (t.ca[i] ? 2.0 : 1.0);
  double bound = ba;
  if (t.b[i] >= 0) bound = ba * (absmax[t.b[i]] * (t.cb[i] ? 2.0 : 1.0));
  return fx_exponent(bound);
}

__device__ __forceinline__ double term_value(const AccTerms &t, int i, int s, const P6 &p) {
  double u = p6_get(p, t.a[i]);
  if (t.ca[i]) u = u - t.ca[i][s];
  if (t.b[i] >= 0) {
    double w = p6_get(p, t.b[i]);
    if (t.cb[i]) w = w - t.cb[i][s];
    u = u * w;
  }
  return u;
}

// True if u*2^e fits the fixed-point range; otherwise the value goes to the
// FP64 fallback accumulator (non-finite values, or a bound that was violated).
__device__ __forceinline__ bool fx_fits(double u, int e) {
  const double q = ldexp(u, e);
  return fabs(q) < 4611686018427387904.0;  // 2^62
}

__device__ __forceinline__ void acc_add_value(const AccLds &L, int S, int i, int s, double u,
                                              int e, double *fallback) {
  if (fx_fits(u, e)) {
    unsigned long long hi, lo;
    fx_split(u, e, hi, lo);
    atomicAdd(&L.hi[(size_t)i * S + s], hi);
    atomicAdd(&L.lo[(size_t)i * S + s], lo);
  } else {
    atomicAdd(&fallback[(size_t)i * S + s], u);
  }
}

__device__ __forceinline__ void acc_add_lds(const AccLds &L, int S, const AccTerms &t,
                                            const int *e, int s, const P6 &p,
                                            double *fallback) {
  atomicAdd(&L.counts[s], 1u);
  for (int i = 0; i < t.n; ++i) acc_add_value(L, S, i, s, term_value(t, i, s, p), e[i], fallback);
}

// Store this block's LDS histogram as partial number h (every entry, zeros
// included, so the merge needs no prior clearing; the slots of a last,
// partly used tile are never read).
__device__ __forceinline__ void acc_lds_store(const AccLds &L, int S, int T, const AccParts &P,
                                              int h) {
  lds_sync();
  for (int s = threadIdx.x; s < S; s += blockDim.x) P.cnt[acc_cnt_index(P.H, h, s)] = L.counts[s];
  // consecutive threads fill one tile's record [t][ACC_TW] after the other
  const int rec = T * ACC_TW;
  const int nq = (int)(acc_tiled(S) / ACC_TW) * rec;
  for (int q = threadIdx.x; q < nq; q += blockDim.x) {
    const int r = q % rec, i = r / ACC_TW;
    const int s = (q / rec) * ACC_TW + r % ACC_TW;
    if (s < S) {
      ulonglong2 v;
      v.x = L.hi[(size_t)i * S + s];
      v.y = L.lo[(size_t)i * S + s];
      P.hl[acc_hl_index(P.H, T, h, i, s)] = v;
    }
  }
}

// Merge of the H partial histograms for a tile of SL slices (RD_SL unless
// the caller picks another), one block of PHX_BLOCK threads: thread t takes
// slice t % SL and the partials h = t / SL, t / SL + G, ... (G = PHX_BLOCK /
// SL); the G groups of a slice are then combined by lane exchanges within
// each wave and across the waves in LDS.  The sums are integers, so the
// order of the additions does not matter.  With the tiled layout and SL a
// multiple of ACC_TW, the lanes of a wave read whole lines that no other
// block reads.  Threads 0 .. SL-1 receive the totals of their slice.
#define RD_SL 4
#define RD_G (PHX_BLOCK / RD_SL)

struct SliceTotals {
  unsigned int count;
  unsigned long long hi[PHX_MAX_TERMS], lo[PHX_MAX_TERMS];
};

// bid: the tile of SL slices (blockIdx.x of the standalone launch)
template <int T, int SL = RD_SL>
__device__ __forceinline__ bool acc_merge(const AccParts &P, int S, SliceTotals &out, int bid) {
  constexpr int NVAL = 2 * T + 1;
  constexpr int NW = PHX_BLOCK / PHX_WAVE;
  __shared__ unsigned long long wsum[NW][NVAL][SL];
  const int sl = threadIdx.x % SL, g = threadIdx.x / SL;
  const int s = bid * SL + sl;
  // v[0] = count, v[1 + 2i] = hi of term i, v[2 + 2i] = lo of term i
  unsigned long long v[NVAL];
#pragma unroll
  for (int i = 0; i < NVAL; ++i) v[i] = 0ULL;
  if (s < S) {
    for (int h = g; h < P.H; h += (PHX_BLOCK / SL)) {
      v[0] += P.cnt[acc_cnt_index(P.H, h, s)];
#pragma unroll
      for (int i = 0; i < T; ++i) {
        const ulonglong2 q = P.hl[acc_hl_index(P.H, T, h, i, s)];
        v[1 + 2 * i] += q.x;
        v[2 + 2 * i] += q.y;
      }
    }
  }
  // lanes of one slice differ in the bits above log2(SL)
#pragma unroll
  for (int off = SL; off < PHX_WAVE; off <<= 1)
#pragma unroll
    for (int i = 0; i < NVAL; ++i) v[i] += __shfl_xor(v[i], off);
  const int lane = threadIdx.x % PHX_WAVE, wv = threadIdx.x / PHX_WAVE;
  if (lane < SL) {
#pragma unroll
    for (int i = 0; i < NVAL; ++i) wsum[wv][i][lane] = v[i];
  }
  lds_sync();
  if (threadIdx.x >= SL || s >= S) return false;
  unsigned long long tot[NVAL];
#pragma unroll
  for (int i = 0; i < NVAL; ++i) {
    tot[i] = 0ULL;
#pragma unroll
    for (int q = 0; q < NW; ++q) tot[i] += wsum[q][i][sl];
  }
  out.count = (unsigned int)tot[0];
#pragma unroll
  for (int i = 0; i < T; ++i) {
    out.hi[i] = tot[1 + 2 * i];
    out.lo[i] = tot[2 + 2 * i];
  }
  return true;
}
