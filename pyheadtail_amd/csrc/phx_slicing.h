// Slicing building blocks shared by the standalone slicing calls and the
// fused engine: the z-moment block that fixes the cuts, and the per-slice
// fixed-point accumulators.
#pragma once

#include "phx_device.h"

size_t phx_max_lds_bytes();

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

__device__ __forceinline__ void zblock_accumulate(double (&v)[ZB_NV], const P6 &p,
                                                  double shift) {
  const double dz = p.z - shift;
  v[0] = v[0] + p.z;
  v[1] = v[1] + dz;
  v[2] = v[2] + dz * dz;
  v[3] = fmin(v[3], p.z);
  v[4] = fmax(v[4], p.z);
  const double a[6] = {fabs(p.x), fabs(p.xp), fabs(p.y), fabs(p.yp), fabs(p.z), fabs(p.dp)};
#pragma unroll
  for (int i = 0; i < 6; ++i)
    if (isfinite(a[i])) v[5 + i] = fmax(v[5 + i], a[i]);
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
  __syncthreads();
  for (int i = threadIdx.x; i <= cfg.n_slices; i += blockDim.x)
    z_bins[i] = arange_value(sh_tail, sh_width, i);
  __syncthreads();
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

struct AccGlobal {
  unsigned int *counts;       // [S]
  unsigned long long *hi;     // [T][S]
  unsigned long long *lo;     // [T][S]
  double *fallback;           // [T][S] non-finite or out-of-scale values
};

struct AccLds {
  unsigned long long *hi;
  unsigned long long *lo;
  unsigned int *counts;
};

// Workspace: [tickets 256 B][chunk partials][counts S u32, padded][hi][lo][fallback]
inline size_t acc_partials_bytes(int64_t n) {
  const int64_t nc = n <= 0 ? 0 : (n + PHX_CHUNK - 1) / PHX_CHUNK;
  return (size_t)nc * 64 * sizeof(double);
}

inline size_t acc_global_bytes(int S, int T) {
  const size_t cb = ((size_t)S * 4 + 255) / 256 * 256;
  return cb + (size_t)T * S * 24;   // multiple of 8 bytes
}

inline AccGlobal acc_global_layout(void *ws, int64_t n, int S, int T) {
  char *p = reinterpret_cast<char *>(ws) + 256 + acc_partials_bytes(n);
  AccGlobal g;
  g.counts = reinterpret_cast<unsigned int *>(p);
  p += ((size_t)S * 4 + 255) / 256 * 256;
  g.hi = reinterpret_cast<unsigned long long *>(p);
  p += (size_t)T * S * 8;
  g.lo = reinterpret_cast<unsigned long long *>(p);
  p += (size_t)T * S * 8;
  g.fallback = reinterpret_cast<double *>(p);
  return g;
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
  __syncthreads();
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

__device__ __forceinline__ void acc_add_lds(const AccLds &L, int S, const AccTerms &t,
                                            const int *e, int s, const P6 &p,
                                            const AccGlobal &g) {
  atomicAdd(&L.counts[s], 1u);
  for (int i = 0; i < t.n; ++i) {
    const double u = term_value(t, i, s, p);
    if (fx_fits(u, e[i])) {
      unsigned long long hi, lo;
      fx_split(u, e[i], hi, lo);
      atomicAdd(&L.hi[(size_t)i * S + s], hi);
      atomicAdd(&L.lo[(size_t)i * S + s], lo);
    } else {
      atomicAdd(&g.fallback[(size_t)i * S + s], u);
    }
  }
}

__device__ __forceinline__ void acc_add_global(const AccGlobal &g, int S, const AccTerms &t,
                                               const int *e, int s, const P6 &p) {
  atomicAdd(&g.counts[s], 1u);
  for (int i = 0; i < t.n; ++i) {
    const double u = term_value(t, i, s, p);
    if (fx_fits(u, e[i])) {
      unsigned long long hi, lo;
      fx_split(u, e[i], hi, lo);
      atomicAdd(&g.hi[(size_t)i * S + s], hi);
      atomicAdd(&g.lo[(size_t)i * S + s], lo);
    } else {
      atomicAdd(&g.fallback[(size_t)i * S + s], u);
    }
  }
}

__device__ __forceinline__ void acc_lds_flush(const AccLds &L, int S, int T,
                                              const AccGlobal &g) {
  __syncthreads();
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const unsigned int c = L.counts[s];
    if (c == 0u) continue;
    atomicAdd(&g.counts[s], c);
    for (int i = 0; i < T; ++i) {
      const unsigned long long h = L.hi[(size_t)i * S + s], l = L.lo[(size_t)i * S + s];
      if (h) atomicAdd(&g.hi[(size_t)i * S + s], h);
      if (l) atomicAdd(&g.lo[(size_t)i * S + s], l);
    }
  }
}

// Last block: exact sums -> outputs, then re-zero the global accumulators.
__device__ __forceinline__ void acc_finalize(const AccGlobal &g, int S, const AccTerms &t,
                                             const int *e, int32_t *counts, double *sums,
                                             double *means) {
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    const unsigned int c = g.counts[s];
    if (counts) counts[s] = (int32_t)c;
    for (int i = 0; i < t.n; ++i) {
      const size_t k = (size_t)i * S + s;
      const double v = fx_value(g.hi[k], g.lo[k], e[i]) + g.fallback[k];
      if (sums) sums[k] = v;
      if (means) means[k] = c ? v / (double)c : 0.0;
      g.hi[k] = 0ULL;
      g.lo[k] = 0ULL;
      g.fallback[k] = 0.0;
    }
    g.counts[s] = 0u;
  }
}
