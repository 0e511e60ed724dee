// Wake-field C entry points: history push, FP64 convolution, kick.
#include "phx_device.h"
#include "phx_host.h"
#include "phx_slicing.h"
#include "phx_wake.h"

namespace {

__global__ void k_hist_push(phx_wake_hist *h, const double *times, const int32_t *counts,
                            const double *mean_x, const double *mean_y, double age_inc,
                            double beta) {
  hist_push_device(h, times, counts, mean_x, mean_y, age_inc, beta);
}

__global__ void __launch_bounds__(WK_THREADS)
    k_wake_conv(const phx_wake_hist *h, WakeKicks kk, int S, double *part, int T, double *out,
                int gmax) {
  wake_conv_block(h, kk, S, part, T, out, gmax, blockIdx.x, blockIdx.y, blockIdx.z, gridDim.y);
}

__global__ void k_wake_rw_desc(const phx_wake_hist *h, WakeKicks kk, int nk, int T, RwKick *desc) {
  rw_desc_build(h, kk, nk, T, desc);
}

template <int KM>
__global__ void __launch_bounds__(rw_threads<KM>())
    k_wake_rw(const phx_wake_hist *h, const RwKick *desc, int nk, int S, double *part, int T,
              double *out) {
  wake_rw_block<KM>(h, desc, nk, S, part, T, out, blockIdx.x);
}

__global__ void k_wake_conv_reduce(const phx_wake_hist *h, WakeKicks kk, int n_kicks, int S,
                                   const double *part, int T, double *out) {
  wake_conv_reduce(h, kk, n_kicks, S, part, T, out, blockIdx.x);
}

__global__ void __launch_bounds__(PHX_BLOCK)
    k_wake_apply(Coords6 co, double *xp, double *yp, double *dp, int64_t n, const int32_t *idx,
                 const phx_slice_state *state, WakeKicks kd, int n_kicks,
                 int S, const double *__restrict__ kicks) {
  const double tail = state->tail, width = state->width;
  const int64_t base = (int64_t)blockIdx.x * PHX_CHUNK + threadIdx.x;
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j >= n) continue;
    int32_t s;
    if (idx) {
      s = idx[j];
    } else {
      const double f = slice_floor(co.c[4][j], tail, width);
      s = (f >= 0.0 && f < (double)S) ? (int32_t)f : -1;
    }
    if (s < 0 || s >= S) continue;
    P6 p;
    p.x = co.c[0] ? co.c[0][j] : 0.0;
    p.y = co.c[2] ? co.c[2][j] : 0.0;
    p.xp = xp ? xp[j] : 0.0;
    p.yp = yp ? yp[j] : 0.0;
    p.dp = dp ? dp[j] : 0.0;
    p.z = 0.0;
    apply_wake_kicks(p, s, S, kd.k, n_kicks, kicks);
    if (xp) xp[j] = p.xp;
    if (yp) yp[j] = p.yp;
    if (dp) dp[j] = p.dp;
  }
}

}  // namespace

extern "C" int phx_wake_hist_push(phx_wake_hist *hist, const double *times,
                                  const int32_t *counts, const double *mean_x,
                                  const double *mean_y, double age_increment, double beta,
                                  phx_stream_t stream) {
  if (!hist || !times || !counts) return phx::fail(PHX_E_INVALID, "phx_wake_hist_push: null pointer");
  k_hist_push<<<1, PHX_BLOCK, 0, phx::as_stream(stream)>>>(hist, times, counts, mean_x, mean_y,
                                                           age_increment, beta);
  PHX_CHECK_LAUNCH("phx_wake_hist_push");
  return PHX_OK;
}

static int check_kicks(const phx_wake_kick *kicks, int32_t n_kicks, const char *who) {
  if (!kicks || n_kicks < 1) return phx::fail(PHX_E_INVALID, "%s: no kicks", who);
  if (n_kicks > PHX_MAX_KICKS) return phx::fail(PHX_E_LIMIT, "%s: n_kicks=%d > %d", who, n_kicks, PHX_MAX_KICKS);
  for (int i = 0; i < n_kicks; ++i) {
    const phx_wake_kick &k = kicks[i];
    if (k.fn < PHX_WAKE_RES_TRANSVERSE || k.fn > PHX_WAKE_TABLE)
      return phx::fail(PHX_E_INVALID, "%s: kick %d bad wake function %d", who, i, k.fn);
    if (k.moment < 0 || k.moment > 2 || k.plane < 0 || k.plane > 2)
      return phx::fail(PHX_E_INVALID, "%s: kick %d bad moment/plane", who, i);
    if (k.multiply != -1 && k.multiply != 0 && k.multiply != 2)
      return phx::fail(PHX_E_INVALID, "%s: kick %d bad multiply", who, i);
    if (k.n_turns_wake < 1) return phx::fail(PHX_E_INVALID, "%s: kick %d n_turns_wake < 1", who, i);
    if (k.fn == PHX_WAKE_TABLE &&
        (!k.table_t || !k.table_w || k.table_n < 2 || k.table_kind < PHX_TABLE_TRANS_UR ||
         k.table_kind > PHX_TABLE_LONG))
      return phx::fail(PHX_E_INVALID, "%s: kick %d bad wake table", who, i);
    if (k.precision < PHX_PREC_FP64 || k.precision > PHX_PREC_MFMA_BF16X6)
      return phx::fail(PHX_E_INVALID, "%s: kick %d bad precision %d", who, i, k.precision);
  }
  return PHX_OK;
}

// The partial sums, then the resistive-wall kick description.
size_t phx_wake_conv_ws_bytes(int32_t n_kicks, int32_t n_turns, int32_t n_slices) {
  return (size_t)n_kicks * conv_rows(n_slices, n_turns) * n_slices * sizeof(double) +
         rw_desc_bytes(n_turns) + 256;
}

int phx_wake_convolve_ws(const phx_wake_hist *hist, const phx_wake_kick *kicks, int32_t n_kicks,
                         int32_t n_slices, int32_t max_turns, double *kicks_out, void *ws,
                         hipStream_t s) {
  const int tiles = (n_slices + WK_TILE - 1) / WK_TILE;
  double *part = reinterpret_cast<double *>(ws);
  WakeKicks kk{};
  int T = 1;
  for (int i = 0; i < n_kicks; ++i) {
    kk.k[i] = kicks[i];
    const int tk = kicks[i].n_turns_wake < max_turns ? kicks[i].n_turns_wake : max_turns;
    T = tk > T ? tk : T;
  }
  if (wake_conv_lds_bytes(n_slices, 1) + WK_STATIC_LDS > phx_max_lds_bytes())
    return phx::fail(PHX_E_LIMIT, "phx_wake_convolve: n_slices=%d too large for LDS", n_slices);
  const int gmax = wake_conv_gmax(n_slices, phx_max_lds_bytes());
  bool general = false, rw_valu = false, rw_mfma = false;
  for (int i = 0; i < n_kicks; ++i) {
    if (!is_rw_dot(kicks[i])) general = true;
    else if (rw_is_mfma(kicks[i].precision)) rw_mfma = true;
    else rw_valu = true;
  }
  const bool rw = rw_valu || rw_mfma;
  if (rw && wake_rw_lds_bytes(n_slices) + RW_STATIC_LDS > phx_max_lds_bytes())
    return phx::fail(PHX_E_LIMIT, "phx_wake_convolve: n_slices=%d too large for LDS", n_slices);
  if (general) {
    k_wake_conv<<<dim3(tiles, n_kicks, T), WK_THREADS, wake_conv_lds_bytes(n_slices, gmax), s>>>(
        hist, kk, n_slices, part, T, kicks_out, gmax);
    PHX_CHECK_LAUNCH("phx_wake_convolve");
  }
  RwKick *desc = reinterpret_cast<RwKick *>(part + (size_t)n_kicks * conv_rows(n_slices, T) * n_slices);
  if (rw) {
    k_wake_rw_desc<<<1, 256, 0, s>>>(hist, kk, n_kicks, T, desc);
    PHX_CHECK_LAUNCH("phx_wake_convolve: resistive-wall kicks");
  }
  if (rw_valu) {
    k_wake_rw<0><<<dim3((n_slices + RW_TILE - 1) / RW_TILE * conv_rows(n_slices, T)), rw_threads<0>(),
                   wake_rw_lds_bytes(n_slices), s>>>(hist, desc, n_kicks, n_slices, part, T,
                                                     kicks_out);
    PHX_CHECK_LAUNCH("phx_wake_convolve: resistive wall");
  }
  if (rw_mfma) {
    k_wake_rw<1><<<dim3((n_slices + RW_TILE - 1) / RW_TILE * conv_rows(n_slices, T)), rw_threads<1>(),
                   wake_rw_lds_bytes(n_slices), s>>>(hist, desc, n_kicks, n_slices, part, T,
                                                     kicks_out);
    PHX_CHECK_LAUNCH("phx_wake_convolve: resistive wall (matrix cores)");
  }
  if (conv_reduces(n_slices, T, rw)) {
    const int tot = n_kicks * n_slices;
    k_wake_conv_reduce<<<dim3((tot * RED_LANES + 255) / 256), 256, 0, s>>>(hist, kk, n_kicks, n_slices, part,
                                                              T, kicks_out);
    PHX_CHECK_LAUNCH("phx_wake_convolve: reduce");
  }
  return PHX_OK;
}

extern "C" int phx_wake_convolve(const phx_wake_hist *hist, const phx_wake_kick *kicks,
                                  int32_t n_kicks, int32_t n_slices, int32_t max_turns,
                                  double *kicks_out, void *ws, phx_stream_t stream) {
  if (!hist || !kicks_out || !ws) return phx::fail(PHX_E_INVALID, "phx_wake_convolve: null pointer");
  if (n_slices < 1) return phx::fail(PHX_E_INVALID, "phx_wake_convolve: n_slices < 1");
  if (max_turns < 1) return phx::fail(PHX_E_INVALID, "phx_wake_convolve: max_turns < 1");
  int rc = check_kicks(kicks, n_kicks, "phx_wake_convolve");
  if (rc) return rc;
  return phx_wake_convolve_ws(hist, kicks, n_kicks, n_slices, max_turns, kicks_out, ws,
                              phx::as_stream(stream));
}

extern "C" int64_t phx_wake_workspace_bytes(int32_t n_kicks, int32_t max_turns,
                                            int32_t n_slices) {
  return (int64_t)phx_wake_conv_ws_bytes(n_kicks, max_turns, n_slices);
}

extern "C" int phx_wake_apply(double *const coords[6], int64_t n, const int32_t *idx,
                              const phx_slice_state *state, const phx_wake_kick *kicks,
                              int32_t n_kicks, int32_t n_slices, const double *kicks_in,
                              phx_stream_t stream) {
  if (!coords || !state || !kicks_in) return phx::fail(PHX_E_INVALID, "phx_wake_apply: null pointer");
  int rc = check_kicks(kicks, n_kicks, "phx_wake_apply");
  if (rc) return rc;
  if (!idx && !coords[4]) return phx::fail(PHX_E_INVALID, "phx_wake_apply: need idx or z");
  WakeKicks kd{};
  for (int i = 0; i < n_kicks; ++i) {
    kd.k[i] = kicks[i];
    const int pl = kicks[i].plane;
    if (!coords[pl == PHX_PLANE_XP ? 1 : (pl == PHX_PLANE_YP ? 3 : 5)])
      return phx::fail(PHX_E_INVALID, "phx_wake_apply: kicked coordinate is null");
    if ((kicks[i].multiply == 0 && !coords[0]) || (kicks[i].multiply == 2 && !coords[2]))
      return phx::fail(PHX_E_INVALID, "phx_wake_apply: multiplying coordinate is null");
  }
  if (n <= 0) return PHX_OK;
  // load and store only the coordinates that the kicks use
  bool t_xp = false, t_yp = false, t_dp = false, u_x = false, u_y = false;
  for (int i = 0; i < n_kicks; ++i) {
    t_xp |= kicks[i].plane == PHX_PLANE_XP;
    t_yp |= kicks[i].plane == PHX_PLANE_YP;
    t_dp |= kicks[i].plane == PHX_PLANE_DP;
    u_x |= kicks[i].multiply == 0;
    u_y |= kicks[i].multiply == 2;
  }
  Coords6 co{};
  co.c[0] = u_x ? coords[0] : nullptr;
  co.c[2] = u_y ? coords[2] : nullptr;
  co.c[4] = idx ? nullptr : coords[4];
  k_wake_apply<<<dim3((unsigned)phx::n_chunks(n)), PHX_BLOCK, 0, phx::as_stream(stream)>>>(
      co, t_xp ? coords[1] : nullptr, t_yp ? coords[3] : nullptr, t_dp ? coords[5] : nullptr, n,
      idx, state, kd, n_kicks, n_slices, kicks_in);
  PHX_CHECK_LAUNCH("phx_wake_apply");
  return PHX_OK;
}
