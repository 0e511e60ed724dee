// The uniform-charge slicer's device part (particles/slicing.py:636-682):
// particles sorted by z (rocPRIM radix sort: stable, where np.argsort's
// quicksort leaves the order of equal z unspecified), the counts below
// the cuts, and the slice index of every particle from the first sorted
// position of each slice.
#include <rocprim/device/device_radix_sort.hpp>

#include "phx_device.h"
#include "phx_host.h"

namespace {

__global__ void k_iota_zero(int32_t *p, int64_t n, unsigned long long *cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = (int32_t)i;
  if (i < 2) cnt[i] = 0ULL;
}

// cnt[0] = #(z < tail), cnt[1] = #(z < head): np.searchsorted(z_sorted, cut)
__global__ void __launch_bounds__(PHX_BLOCK)
    k_count_below(const double *__restrict__ z, int64_t n, const phx_slice_state *state,
                  unsigned long long *cnt) {
  __shared__ unsigned int part[2][PHX_BLOCK / 64];
  const double tail = state->tail, head = state->head;
  unsigned int a = 0, b = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = z[i];
    a += v < tail;
    b += v < head;
  }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = a;
    part[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sa = 0, sb = 0;
    for (int w = 0; w < PHX_BLOCK / 64; ++w) {
      sa += part[0][w];
      sb += part[1][w];
    }
    atomicAdd(&cnt[0], sa);
    atomicAdd(&cnt[1], sb);
  }
}

// Sorted position r belongs to slice i for first[i] <= r < first[i+1]
// (i < S), else to none (-1); written back through the permutation.  Thread
// i < S - 1 also writes the bin edge between slices i and i + 1, the mean
// of the two sorted z around it (a negative position wraps, as NumPy's
// indexing does); the outer edges are the cuts.
__global__ void __launch_bounds__(PHX_BLOCK)
    k_ucs_assign(const double *__restrict__ zs, const int32_t *__restrict__ perm, int64_t n,
                 int S, const int32_t *__restrict__ first, const phx_slice_state *state,
                 int32_t *__restrict__ idx, double *__restrict__ z_bins) {
  const int64_t r0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r0 == 0) {
    z_bins[0] = state->tail;
    z_bins[S] = state->head;
  }
  if (r0 < S - 1) {
    const int64_t hi = first[r0 + 1];
    if (hi < 0 || hi >= n) {
      z_bins[r0 + 1] = NAN;  // no particle after the edge (NumPy: IndexError)
    } else {
      const int64_t lo = hi > 0 ? hi - 1 : n - 1;
      z_bins[r0 + 1] = (zs[lo] + zs[hi]) / 2;
    }
  }
  for (int64_t r = r0; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    int32_t s = -1;
    if (r >= first[0] && r < first[S]) {
      int lo = 0, hi = S;  // first[lo] <= r < first[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= r) lo = mid;
        else hi = mid;
      }
      s = lo;
    }
    idx[perm[r]] = s;
  }
}

size_t radix_temp_bytes(int64_t n) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs(nullptr, bytes, (const double *)nullptr, (double *)nullptr,
                                  (const int32_t *)nullptr, (int32_t *)nullptr, (size_t)n, 0,
                                  64, (hipStream_t)0);
  return bytes;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" int64_t phx_ucs_workspace_bytes(int64_t n) {
  if (n < 1) return -1;
  return (int64_t)(256 + align256((size_t)n * sizeof(int32_t)) + align256(radix_temp_bytes(n)));
}

extern "C" int phx_ucs_sort(const double *z, int64_t n, const phx_slice_state *state,
                            double *z_sorted, int32_t *perm, uint64_t *below, void *ws,
                            phx_stream_t stream) {
  if (!z || !state || !z_sorted || !perm || !below || !ws)
    return phx::fail(PHX_E_INVALID, "phx_ucs_sort: null pointer");
  if (n < 1) return phx::fail(PHX_E_INVALID, "phx_ucs_sort: n < 1");
  if (n > 0x7fffffffLL) return phx::fail(PHX_E_LIMIT, "phx_ucs_sort: n too large");
  hipStream_t s = phx::as_stream(stream);
  char *w = reinterpret_cast<char *>(ws) + 256;
  int32_t *iota = reinterpret_cast<int32_t *>(w);
  void *tmp = w + align256((size_t)n * sizeof(int32_t));
  size_t tmp_bytes = radix_temp_bytes(n);
  unsigned long long *cnt = reinterpret_cast<unsigned long long *>(below);
  k_iota_zero<<<dim3((unsigned)((n + 255) / 256)), 256, 0, s>>>(iota, n, cnt);
  PHX_CHECK_LAUNCH("phx_ucs_sort: iota");
  const hipError_t e = rocprim::radix_sort_pairs(tmp, tmp_bytes, z, z_sorted, iota, perm,
                                                 (size_t)n, 0, 64, s);
  if (e != hipSuccess) return phx::fail((int)e, "phx_ucs_sort: radix sort: %s", hipGetErrorString(e));
  const int64_t g = phx::n_chunks(n);
  k_count_below<<<dim3((unsigned)(g < 1024 ? g : 1024)), PHX_BLOCK, 0, s>>>(z, n, state, cnt);
  PHX_CHECK_LAUNCH("phx_ucs_sort: count");
  return PHX_OK;
}

extern "C" int phx_ucs_assign(const double *z_sorted, const int32_t *perm, int64_t n,
                              int32_t n_slices, const int32_t *first, const phx_slice_state *state,
                              int32_t *idx, double *z_bins, phx_stream_t stream) {
  if (!z_sorted || !perm || !first || !state || !idx || !z_bins)
    return phx::fail(PHX_E_INVALID, "phx_ucs_assign: null pointer");
  if (n < 1 || n_slices < 1) return phx::fail(PHX_E_INVALID, "phx_ucs_assign: n or n_slices < 1");
  const int64_t blocks = ((n > n_slices ? n : (int64_t)n_slices) + PHX_BLOCK - 1) / PHX_BLOCK;
  k_ucs_assign<<<dim3((unsigned)(blocks < 4096 ? blocks : 4096)), PHX_BLOCK, 0,
                 phx::as_stream(stream)>>>(z_sorted, perm, n, n_slices, first, state, idx,
                                           z_bins);
  PHX_CHECK_LAUNCH("phx_ucs_assign");
  return PHX_OK;
}
