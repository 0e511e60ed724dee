// Bunch and per-slice statistics, and uniform-bin slicing.
//
// Reference behaviour restated here:
//   np.mean (general/pmath.py:183), cobra_functions/stats.pyx:cov (:17-45)
//   Slicer.get_long_cuts (particles/slicing.py:395-413)
//   UniformBinSlicer.compute_sliceset_kwargs (particles/slicing.py:599-615)
//   count_macroparticles_per_slice / mean_per_slice / std_per_slice /
//   cov_per_slice / emittance_per_slice (cobra_functions/stats.pyx:192-425)
#include "phx_device.h"
#include "phx_host.h"
#include "phx_slicing.h"

// ---------------------------------------------------------------------------
// Bunch sums of K arrays: slots [sum a (K)] [sum a-a0 (K)] [pairs (P)]
// [min (K)] [max (K)] [max|a| finite (K)].
// ---------------------------------------------------------------------------
struct Arrays6 {
  const double *a[6];
};

template <int K>
struct BunchLayout {
  static constexpr int P = K * (K + 1) / 2;
  static constexpr int NV = 5 * K + P;
  static constexpr int MIN0 = 2 * K + P;
  static constexpr unsigned long long MINM = ((1ULL << K) - 1ULL) << MIN0;
  static constexpr unsigned long long MAXM = ((1ULL << (2 * K)) - 1ULL) << (MIN0 + K);
};

template <int K>
__global__ void __launch_bounds__(PHX_BLOCK)
    k_bunch_sums(Arrays6 arr, int64_t n, double *partials) {
  using L = BunchLayout<K>;
  __shared__ double lds[BR_LDS(L::NV)];
  double shift[K];
#pragma unroll
  for (int i = 0; i < K; ++i) shift[i] = arr.a[i][0];
  double v[L::NV];
#pragma unroll
  for (int i = 0; i < L::NV; ++i) v[i] = red_ident<L::MINM, L::MAXM>(i);
  const int64_t ng = phx_n_groups(n);
  for (int64_t gi = blockIdx.x; gi < ng; gi += gridDim.x) {
  const int64_t base = gi * PHX_CHUNK;
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK + threadIdx.x;
    if (j < n) {
      double u[K];
#pragma unroll
      for (int i = 0; i < K; ++i) u[i] = arr.a[i][j];
      int p = 0;
#pragma unroll
      for (int i = 0; i < K; ++i) {
        v[i] = v[i] + u[i];
        const double di = u[i] - shift[i];
        v[K + i] = v[K + i] + di;
#pragma unroll
        for (int jj = i; jj < K; ++jj) {
          v[2 * K + p] = v[2 * K + p] + di * (u[jj] - shift[jj]);
          ++p;
        }
        v[L::MIN0 + i] = fmin(v[L::MIN0 + i], u[i]);
        v[L::MIN0 + K + i] = fmax(v[L::MIN0 + K + i], u[i]);
        const double au = fabs(u[i]);
        if (isfinite(au)) v[L::MIN0 + 2 * K + i] = fmax(v[L::MIN0 + 2 * K + i], au);
      }
    }
  }
  }
  const double r = block_reduce_slots<L::NV, L::MINM, L::MAXM>(v, lds);
  if (threadIdx.x < L::NV) partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = r;
}

// One block folds the chunk partials in a fixed order.
template <int K>
__global__ void __launch_bounds__(PHX_BLOCK)
    k_bunch_sums_fin(const double *partials, int64_t n_chunks, double *out) {
  using L = BunchLayout<K>;
  __shared__ double res[L::NV];
  fold_partials<L::MINM, L::MAXM>(partials, n_chunks, 0, L::NV, res);
  if (threadIdx.x < L::NV) out[threadIdx.x] = res[threadIdx.x];
}

extern "C" int64_t phx_bunch_sums_len(int32_t k) {
  if (k < 1 || k > 6) return -1;
  return 5 * k + k * (k + 1) / 2;
}

extern "C" int phx_bunch_sums(const double *const *arrays, int32_t k, int64_t n, double *out,
                              void *ws, phx_stream_t stream) {
  if (k < 1 || k > 6) return phx::fail(PHX_E_INVALID, "phx_bunch_sums: k=%d not in 1..6", k);
  if (n < 1) return phx::fail(PHX_E_INVALID, "phx_bunch_sums: n=%lld < 1", (long long)n);
  if (!arrays || !out || !ws) return phx::fail(PHX_E_INVALID, "phx_bunch_sums: null pointer");
  Arrays6 a{};
  for (int i = 0; i < k; ++i) {
    if (!arrays[i]) return phx::fail(PHX_E_INVALID, "phx_bunch_sums: array %d is null", i);
    a.a[i] = arrays[i];
  }
  double *partials = reinterpret_cast<double *>(reinterpret_cast<char *>(ws) + phx::WsLayout::kTicketBytes);
  const int64_t nc = phx_pass_grid(n);
  const dim3 grid((unsigned)nc);
  hipStream_t s = phx::as_stream(stream);
#define PHX_BS(KK)                                                              \
  case KK:                                                                      \
    k_bunch_sums<KK><<<grid, PHX_BLOCK, 0, s>>>(a, n, partials);                \
    k_bunch_sums_fin<KK><<<1, PHX_BLOCK, 0, s>>>(partials, nc, out);            \
    break;
  switch (k) {
    PHX_BS(1)
    PHX_BS(2)
    PHX_BS(3)
    PHX_BS(4)
    PHX_BS(5)
    default:
      k_bunch_sums<6><<<grid, PHX_BLOCK, 0, s>>>(a, n, partials);
      k_bunch_sums_fin<6><<<1, PHX_BLOCK, 0, s>>>(partials, nc, out);
      break;
  }
#undef PHX_BS
  PHX_CHECK_LAUNCH("phx_bunch_sums");
  return PHX_OK;
}

// ---------------------------------------------------------------------------
// Slicing cuts: z sums of the current coordinates and max|u| of all six.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(PHX_BLOCK)
    k_slice_cuts(Coords6 co, int64_t n, double *partials) {
  __shared__ double lds[BR_LDS(ZB_NV)];
  const double shift = co.c[4][0];
  double v[ZB_NV];
#pragma unroll
  for (int i = 0; i < ZB_NV; ++i) v[i] = red_ident<ZB_MINM, ZB_MAXM>(i);
  const int64_t ng = phx_n_groups(n);
  for (int64_t gi = blockIdx.x; gi < ng; gi += gridDim.x) {
  const int64_t base = gi * PHX_CHUNK;
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK + threadIdx.x;
    if (j < n) {
      P6 p;
      p.x = co.c[0] ? co.c[0][j] : 0.0;
      p.xp = co.c[1] ? co.c[1][j] : 0.0;
      p.y = co.c[2] ? co.c[2][j] : 0.0;
      p.yp = co.c[3] ? co.c[3][j] : 0.0;
      p.z = co.c[4][j];
      p.dp = co.c[5] ? co.c[5][j] : 0.0;
      zblock_accumulate(v, p, shift);
    }
  }
  }
  const double r = block_reduce_slots<ZB_NV, ZB_MINM, ZB_MAXM>(v, lds);
  if (threadIdx.x < ZB_NV) partials[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = r;
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_slice_cuts_fin(const double *partials, int64_t n_chunks, phx_slicer cfg, int64_t n,
                     phx_slice_state *state, double *z_bins) {
  __shared__ double res[ZB_NV];
  fold_partials<ZB_MINM, ZB_MAXM>(partials, n_chunks, 0, ZB_NV, res);
  zblock_finalize(res, cfg, n, state, z_bins, -1);
}

extern "C" int phx_slice_cuts(const double *const coords[6], int64_t n, const phx_slicer *cfg,
                              phx_slice_state *state, double *z_bins, void *ws,
                              phx_stream_t stream) {
  if (!coords || !coords[4] || !cfg || !state || !z_bins || !ws)
    return phx::fail(PHX_E_INVALID, "phx_slice_cuts: null pointer");
  if (n < 1) return phx::fail(PHX_E_INVALID, "phx_slice_cuts: n=%lld < 1", (long long)n);
  if (cfg->n_slices < 1)
    return phx::fail(PHX_E_INVALID, "phx_slice_cuts: n_slices=%d < 1", cfg->n_slices);
  if (cfg->cuts_mode < PHX_CUTS_FIXED || cfg->cuts_mode > PHX_CUTS_MINMAX)
    return phx::fail(PHX_E_INVALID, "phx_slice_cuts: bad cuts_mode %d", cfg->cuts_mode);
  Coords6 co{};
  for (int i = 0; i < 6; ++i) co.c[i] = coords[i];
  double *partials = reinterpret_cast<double *>(reinterpret_cast<char *>(ws) + phx::WsLayout::kTicketBytes);
  const int64_t nc = phx_pass_grid(n);
  hipStream_t s = phx::as_stream(stream);
  k_slice_cuts<<<dim3((unsigned)nc), PHX_BLOCK, 0, s>>>(co, n, partials);
  k_slice_cuts_fin<<<1, PHX_BLOCK, 0, s>>>(partials, nc, *cfg, n, state, z_bins);
  PHX_CHECK_LAUNCH("phx_slice_cuts");
  return PHX_OK;
}

// ---------------------------------------------------------------------------
// Per-slice fixed-point accumulation.
// ---------------------------------------------------------------------------
__global__ void k_fill_zero(unsigned long long *p, size_t words) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ULL;
}

#define ACC_BLOCKS 256

__global__ void __launch_bounds__(PHX_BLOCK)
    k_slice_accumulate(Coords6 co, int64_t n, const phx_slice_state *state, int S, AccTerms terms,
                       const int32_t *idx_in, int32_t *idx_out, AccParts P) {
  extern __shared__ unsigned long long acc_lds[];
  const double tail = state->tail, width = state->width;
  int e[PHX_MAX_TERMS];
  for (int t = 0; t < terms.n; ++t) e[t] = term_exponent(terms, t, state->absmax);
  AccLds L = acc_lds_layout(acc_lds, S, terms.n);
  acc_lds_zero(L, S, terms.n);
  const int64_t n_chunks = (n + PHX_CHUNK - 1) / PHX_CHUNK;
  for (int64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const int64_t base = c * PHX_CHUNK;
#pragma unroll
    for (int k = 0; k < PHX_PPT; ++k) {
      const int64_t j = base + (int64_t)k * PHX_BLOCK + threadIdx.x;
      if (j < n) {
        const double z = co.c[4] ? co.c[4][j] : 0.0;
        int s;
        if (idx_in) {
          s = idx_in[j];
        } else {
          s = x86_cvt_i32(slice_floor(z, tail, width));
        }
        if (idx_out) idx_out[j] = s;
        if (s >= 0 && s < S) {
          P6 p;
          p.x = co.c[0] ? co.c[0][j] : 0.0;
          p.xp = co.c[1] ? co.c[1][j] : 0.0;
          p.y = co.c[2] ? co.c[2][j] : 0.0;
          p.yp = co.c[3] ? co.c[3][j] : 0.0;
          p.z = z;
          p.dp = co.c[5] ? co.c[5][j] : 0.0;
          acc_add_lds(L, S, terms, e, s, p, P.fallback);
        }
      }
    }
  }
  acc_lds_store(L, S, terms.n, P, blockIdx.x);
}

// Merge the partial histograms: counts, exact sums and means per slice;
// re-zero the FP64 fallback entries.
template <int T>
__global__ void __launch_bounds__(RD_SL * RD_G)
    k_slice_acc_merge(AccParts P, int S, AccTerms terms, const phx_slice_state *state,
                      int32_t *counts, double *sums, double *means) {
  SliceTotals tot;
  if (!acc_merge<T>(P, S, tot, blockIdx.x)) return;
  const int s = blockIdx.x * RD_SL + threadIdx.x % RD_SL;
  if (counts) counts[s] = (int32_t)tot.count;
#pragma unroll
  for (int i = 0; i < T; ++i) {
    const int e = term_exponent(terms, i, state->absmax);
    const size_t k = (size_t)i * S + s;
    const double v = fx_value(tot.hi[i], tot.lo[i], e) + P.fallback[k];
    P.fallback[k] = 0.0;
    if (sums) sums[k] = v;
    if (means) means[k] = tot.count ? v / (double)tot.count : 0.0;
  }
}

extern "C" int phx_slice_accumulate(const double *const coords[6], int64_t n,
                                    const phx_slice_state *state, int32_t n_slices,
                                    const phx_term *terms, int32_t n_terms,
                                    const int32_t *idx_in, int32_t *idx_out, int32_t *counts,
                                    double *sums, double *means, void *ws,
                                    phx_stream_t stream) {
  if (!coords || (!coords[4] && !idx_in) || !state || !counts || !ws)
    return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: null pointer");
  if (n < 1) return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: n=%lld < 1", (long long)n);
  if (n_slices < 1) return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: n_slices < 1");
  if (n_terms < 0 || n_terms > PHX_MAX_TERMS)
    return phx::fail(PHX_E_LIMIT, "phx_slice_accumulate: n_terms=%d not in 0..%d", n_terms,
                     PHX_MAX_TERMS);
  if (n_terms > 0 && (!terms || !sums))
    return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: terms/sums null");
  const size_t lds_bytes = acc_lds_bytes(n_slices, n_terms);
  if (lds_bytes > phx_max_lds_bytes())
    return phx::fail(PHX_E_LIMIT,
                     "phx_slice_accumulate: %d slices x %d terms need %zu B of LDS (max %zu)",
                     n_slices, n_terms, lds_bytes, phx_max_lds_bytes());
  AccTerms t{};
  t.n = n_terms;
  for (int i = 0; i < n_terms; ++i) {
    if (terms[i].a < 0 || terms[i].a > 5 || terms[i].b > 5)
      return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: bad coordinate index in term %d", i);
    if (!coords[terms[i].a] || (terms[i].b >= 0 && !coords[terms[i].b]))
      return phx::fail(PHX_E_INVALID, "phx_slice_accumulate: term %d uses a null coordinate", i);
    t.a[i] = terms[i].a;
    t.b[i] = terms[i].b;
    t.ca[i] = terms[i].center_a;
    t.cb[i] = terms[i].center_b;
  }
  Coords6 co{};
  for (int i = 0; i < 6; ++i) co.c[i] = coords[i];
  const int32_t S = n_slices;
  const int64_t nc = phx::n_chunks(n);
  const int H = (int)(nc < ACC_BLOCKS ? nc : ACC_BLOCKS);
  char *base = reinterpret_cast<char *>(ws) + phx::WsLayout::kTicketBytes + acc_partials_bytes(n);
  AccParts P = acc_parts_layout(base, H, S, n_terms);
  hipStream_t s = phx::as_stream(stream);
  // the FP64 fallback region may hold data of earlier calls with other
  // sizes: clear it (the merge leaves it zero again)
  const size_t fw = (size_t)n_terms * S;
  if (fw) {
    k_fill_zero<<<dim3((unsigned)((fw + 255) / 256 < 1024 ? (fw + 255) / 256 : 1024)), 256, 0,
                  s>>>(reinterpret_cast<unsigned long long *>(P.fallback), fw);
    PHX_CHECK_LAUNCH("phx_slice_accumulate: clear");
  }
  k_slice_accumulate<<<dim3(H), PHX_BLOCK, lds_bytes, s>>>(co, n, state, S, t, idx_in, idx_out,
                                                           P);
  PHX_CHECK_LAUNCH("phx_slice_accumulate");
  const dim3 mg((S + RD_SL - 1) / RD_SL);
  switch (n_terms) {
#define PHX_MG(TT) \
    case TT: k_slice_acc_merge<TT><<<mg, RD_SL * RD_G, 0, s>>>(P, S, t, state, counts, sums, means); break;
    PHX_MG(0) PHX_MG(1) PHX_MG(2) PHX_MG(3) PHX_MG(4) PHX_MG(5) PHX_MG(6) PHX_MG(7)
    default: k_slice_acc_merge<8><<<mg, RD_SL * RD_G, 0, s>>>(P, S, t, state, counts, sums, means);
#undef PHX_MG
  }
  PHX_CHECK_LAUNCH("phx_slice_accumulate: merge");
  return PHX_OK;
}

// ---------------------------------------------------------------------------
// Per-slice std / cov / emittance (cobra_functions/stats.pyx:258-425).
// ---------------------------------------------------------------------------
__global__ void k_slice_finalize(int kind, int S, const int32_t *counts, const double *sums,
                                 int n_sums, double f_pre, double f_mul, double f_div,
                                 double *out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const int32_t nm = counts[s];
  double r;
  if (kind == 0 || kind == 1) {
    double v = sums[s];
    if (nm > 1) v = v / (double)(nm - 1);
    r = kind == 0 ? sqrt(v) : v;
  } else {
    // emittance_per_slice: covariances are each divided by (n-1) when n > 1
    auto cov = [&](int i) {
      double v = sums[(size_t)i * S + s];
      if (nm > 1) v = v / (double)(nm - 1);
      return v;
    };
    if (nm > 1) {
      double s11 = cov(0), s12 = cov(1), s22 = cov(2);
      if (n_sums == 6) {
        const double cdp2 = cov(5);
        if (cdp2 != 0.0) {
          const double cudp = cov(3), cupdp = cov(4);
          s11 = s11 - cudp * cudp / cdp2;
          s12 = s12 - cudp * cupdp / cdp2;
          s22 = s22 - cupdp * cupdp / cdp2;
        }
      }
      r = sqrt(s11 * s22 - s12 * s12);
    } else {
      r = 0.0;
    }
  }
  out[s] = ((f_pre * r) * f_mul) / f_div;
}

extern "C" int phx_slice_finalize(int32_t kind, int32_t n_slices, const int32_t *counts,
                                  const double *sums, int32_t n_sums, double f_pre,
                                  double f_mul, double f_div, double *out,
                                  phx_stream_t stream) {
  if (!counts || !sums || !out) return phx::fail(PHX_E_INVALID, "phx_slice_finalize: null pointer");
  if (n_slices < 1) return phx::fail(PHX_E_INVALID, "phx_slice_finalize: n_slices < 1");
  if (kind < 0 || kind > 2) return phx::fail(PHX_E_INVALID, "phx_slice_finalize: bad kind %d", kind);
  if (kind == 2 && n_sums != 3 && n_sums != 6)
    return phx::fail(PHX_E_INVALID, "phx_slice_finalize: emittance needs 3 or 6 sums");
  const int threads = 256;
  k_slice_finalize<<<dim3((n_slices + threads - 1) / threads), threads, 0,
                     phx::as_stream(stream)>>>(kind, n_slices, counts, sums, n_sums, f_pre,
                                               f_mul, f_div, out);
  PHX_CHECK_LAUNCH("phx_slice_finalize");
  return PHX_OK;
}

__global__ void k_slice_times(const double *zb, int S, double beta_c, double *zc, double *t) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S) return;
  const double c = zb[s] + 0.5 * (zb[s + 1] - zb[s]);
  if (zc) zc[s] = c;
  if (t) t[s] = c / beta_c;
}

extern "C" int phx_slice_times(const double *z_bins, int32_t n_slices, double beta_c,
                               double *z_centers, double *times, phx_stream_t stream) {
  if (!z_bins || n_slices < 1) return phx::fail(PHX_E_INVALID, "phx_slice_times: bad arguments");
  k_slice_times<<<dim3((n_slices + 255) / 256), 256, 0, phx::as_stream(stream)>>>(
      z_bins, n_slices, beta_c, z_centers, times);
  PHX_CHECK_LAUNCH("phx_slice_times");
  return PHX_OK;
}
