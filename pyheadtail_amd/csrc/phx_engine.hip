// Fused one-turn map.  One turn of [local elements..., WakeField] is three
// launches:
//   1. particle pass: apply the wake kick left pending by the previous turn
//      (it needs only z and the previous cuts, both unchanged since), record
//      the bunch moments of that end-of-turn state, run the local elements
//      (longitudinal maps, transverse segments), write the six coordinates
//      once, and reduce the z moments of the new state; the last block turns
//      them into the slicing cuts (96 B of HBM traffic per macroparticle);
//   2. slice histogram: counts and fixed-point sums of x and y per slice
//      (24 B per macroparticle); the last block forms the slice means and
//      pushes them into the multi-turn wake history;
//   3. wake convolution over the history (per-slice work only).
// The kick of turn t is applied in pass 1 of turn t+1, which gives the same
// arithmetic as WakeField.track at the end of turn t
// (impedances/wakes.py:121-145).
#include "phx_device.h"
#include "phx_host.h"
#include "phx_slicing.h"
#include "phx_wake.h"

#define EN_STAT_NV 33
#define EN_NV (ZB_NV + EN_STAT_NV)
#define EN_HIST_BLOCKS 1024

namespace {

struct EngineWs {
  unsigned int *tickets;       // [0] pass, [1] hist, [2] flush
  unsigned int *conv_tickets;  // [PHX_MAX_KICKS * tiles]
  double *partials;            // [chunks][EN_NV]
  AccGlobal acc;               // T = 2 (x, y)
  double *conv_part;           // [n_kicks][T][S]
};

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

__host__ __device__ inline EngineWs engine_ws(void *ws, int64_t n, int S, int n_kicks, int T) {
  EngineWs w;
  char *p = reinterpret_cast<char *>(ws);
  w.tickets = reinterpret_cast<unsigned int *>(p);
  p += 256;
  const int tiles = (S + WK_TILE - 1) / WK_TILE;
  w.conv_tickets = reinterpret_cast<unsigned int *>(p);
  p += ((size_t)PHX_MAX_KICKS * tiles * 4 + 255) / 256 * 256;
  const int64_t nc = (n + PHX_CHUNK - 1) / PHX_CHUNK;
  w.partials = reinterpret_cast<double *>(p);
  p += ((size_t)nc * EN_NV * 8 + 255) / 256 * 256;
  w.acc.counts = reinterpret_cast<unsigned int *>(p);
  p += ((size_t)S * 4 + 255) / 256 * 256;
  w.acc.hi = reinterpret_cast<unsigned long long *>(p);
  p += (size_t)2 * S * 8;
  w.acc.lo = reinterpret_cast<unsigned long long *>(p);
  p += (size_t)2 * S * 8;
  w.acc.fallback = reinterpret_cast<double *>(p);
  p += (size_t)2 * S * 8;
  w.conv_part = reinterpret_cast<double *>(p);
  (void)n_kicks;
  (void)T;
  return w;
}

size_t engine_ws_bytes(int64_t n, int S, int n_kicks, int T) {
  const int tiles = (S + WK_TILE - 1) / WK_TILE;
  const int64_t nc = (n + PHX_CHUNK - 1) / PHX_CHUNK;
  size_t b = 256 + round256((size_t)PHX_MAX_KICKS * tiles * 4) +
             round256((size_t)nc * EN_NV * 8) + round256((size_t)S * 4) + (size_t)6 * S * 8;
  b += (size_t)n_kicks * T * S * 8;
  return b;
}

struct EngineArgs {
  phx_engine e;
  int n_kicks;
  int T;
};

__device__ __forceinline__ P6 load_p(const phx_engine &e, int64_t j) {
  P6 p;
  p.x = e.coords[0][j];
  p.xp = e.coords[1][j];
  p.y = e.coords[2][j];
  p.yp = e.coords[3][j];
  p.z = e.coords[4][j];
  p.dp = e.coords[5][j];
  return p;
}

__device__ __forceinline__ void store_p(const phx_engine &e, int64_t j, const P6 &p) {
  e.coords[0][j] = p.x;
  e.coords[1][j] = p.xp;
  e.coords[2][j] = p.y;
  e.coords[3][j] = p.yp;
  e.coords[4][j] = p.z;
  e.coords[5][j] = p.dp;
}

__device__ __forceinline__ void kick_pending(P6 &p, const phx_turn_program &g, double tail,
                                             double width, int S, const double *kicks) {
  const double f = slice_floor(p.z, tail, width);
  if (f >= 0.0 && f < (double)S)
    apply_wake_kicks(p, (int32_t)f, S, g.kick, g.n_kicks, kicks);
}

// stats slots: [0..6) sum u, [6..12) sum (u-u0), [12..33) sum (u-u0)(v-v0)
__device__ __forceinline__ void stats_accumulate(double *v, const P6 &p, const double *sh) {
  const double u[6] = {p.x, p.xp, p.y, p.yp, p.z, p.dp};
  double d[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    d[i] = u[i] - sh[i];
    v[i] = v[i] + u[i];
    v[6 + i] = v[6 + i] + d[i];
  }
  int q = 12;
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = i; j < 6; ++j) {
      v[q] = v[q] + d[i] * d[j];
      ++q;
    }
}

// MODE 0: a turn; MODE 1: flush (pending kick + final statistics only).
template <int MODE, bool REC, bool WAKE>
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_pass(EngineArgs a) {
  const phx_engine &e = a.e;
  constexpr int NV = (WAKE ? ZB_NV : 0) + (REC ? EN_STAT_NV : 0) + ((!WAKE && !REC) ? 1 : 0);
  constexpr unsigned long long MINM = WAKE ? ZB_MINM : 0ULL;
  constexpr unsigned long long MAXM = WAKE ? ZB_MAXM : 0ULL;
  constexpr int SOFF = WAKE ? ZB_NV : 0;
  __shared__ double lds[NV * (PHX_BLOCK / PHX_WAVE)];
  __shared__ double res[NV];
  __shared__ double sh[8];
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices, a.n_kicks, a.T);
  const int turn = *e.turn_counter;
  const int prog_turn = MODE == 0 ? turn : (turn > 0 ? turn - 1 : 0);
  const phx_turn_program &g = e.programs[e.turn_program[prog_turn]];
  const int pend = *e.pending;
  const double tail = e.state->tail, width = e.state->width;
  const int S = e.n_slices;
  if (threadIdx.x == 0) {
    P6 p = load_p(e, 0);
    if (pend) kick_pending(p, g, tail, width, S, e.kicks);
    sh[0] = p.x; sh[1] = p.xp; sh[2] = p.y; sh[3] = p.yp; sh[4] = p.z; sh[5] = p.dp;
    if (MODE == 0) apply_program(p, g);
    sh[6] = p.z;
  }
  __syncthreads();
  double v[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) v[i] = red_ident<MINM, MAXM>(i);
  double shs[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) shs[i] = sh[i];
  const double zshift = sh[6];
  const int64_t base = (int64_t)blockIdx.x * PHX_CHUNK + threadIdx.x;
  P6 p[PHX_PPT];
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < e.n) p[k] = load_p(e, j);
  }
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int64_t j = base + (int64_t)k * PHX_BLOCK;
    if (j < e.n) {
      if (pend) kick_pending(p[k], g, tail, width, S, e.kicks);
      if (REC) stats_accumulate(&v[SOFF], p[k], shs);
      if (MODE == 0) apply_program(p[k], g);
      if (MODE == 0 || pend) store_p(e, j, p[k]);
      if (WAKE) {
        double zb[ZB_NV];
#pragma unroll
        for (int i = 0; i < ZB_NV; ++i) zb[i] = v[i];
        zblock_accumulate(zb, p[k], zshift);
#pragma unroll
        for (int i = 0; i < ZB_NV; ++i) v[i] = zb[i];
      }
    }
  }
  const double r = block_reduce_slots<NV, MINM, MAXM>(v, lds);
  if (threadIdx.x < NV) w.partials[(int64_t)blockIdx.x * EN_NV + threadIdx.x] = r;
  if (last_block(&w.tickets[MODE == 0 ? 0 : 2], gridDim.x)) {
    // partials of this pass use stride EN_NV; fold them with the NV layout
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = red_ident<MINM, MAXM>(i);
    for (int64_t c = threadIdx.x; c < (int64_t)gridDim.x; c += PHX_BLOCK) {
      const double *pc = w.partials + c * EN_NV;
#pragma unroll
      for (int i = 0; i < NV; ++i) acc[i] = red_slot<MINM, MAXM>(i, acc[i], pc[i]);
    }
    const double rr = block_reduce_slots<NV, MINM, MAXM>(acc, lds);
    if (threadIdx.x < NV) res[threadIdx.x] = rr;
    __syncthreads();
    if (WAKE) zblock_finalize(res, g.slicer, e.n, e.state, e.z_bins, turn);
    if (REC && threadIdx.x < EN_STAT_NV)
      e.stats[(size_t)turn * EN_STAT_NV + threadIdx.x] = res[SOFF + threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
      if (MODE == 0) *e.turn_counter = turn + 1;
      else *e.pending = 0;
    }
  }
}

// Slice histogram of the new state: counts, sum x, sum y per slice.  The
// last block forms the means and pushes the SliceSet data into the history.
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_hist(EngineArgs a) {
  extern __shared__ unsigned long long acc_lds[];
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices, a.n_kicks, a.T);
  const int S = e.n_slices;
  const double tail = e.state->tail, width = e.state->width;
  AccTerms t{};
  t.n = 2;
  t.a[0] = 0; t.b[0] = -1;
  t.a[1] = 2; t.b[1] = -1;
  int ex[2];
  ex[0] = fx_exponent(e.state->absmax[0]);
  ex[1] = fx_exponent(e.state->absmax[2]);
  AccLds L = acc_lds_layout(acc_lds, S, 2);
  acc_lds_zero(L, S, 2);
  const int64_t nc = (e.n + PHX_CHUNK - 1) / PHX_CHUNK;
  for (int64_t c = blockIdx.x; c < nc; c += gridDim.x) {
    const int64_t base = c * PHX_CHUNK + threadIdx.x;
    double zz[PHX_PPT], xx[PHX_PPT], yy[PHX_PPT];
#pragma unroll
    for (int k = 0; k < PHX_PPT; ++k) {
      const int64_t j = base + (int64_t)k * PHX_BLOCK;
      if (j < e.n) {
        zz[k] = e.coords[4][j];
        xx[k] = e.coords[0][j];
        yy[k] = e.coords[2][j];
      }
    }
#pragma unroll
    for (int k = 0; k < PHX_PPT; ++k) {
      const int64_t j = base + (int64_t)k * PHX_BLOCK;
      if (j < e.n) {
        const double f = slice_floor(zz[k], tail, width);
        if (f >= 0.0 && f < (double)S) {
          P6 p;
          p.x = xx[k]; p.y = yy[k]; p.z = zz[k];
          p.xp = p.yp = p.dp = 0.0;
          acc_add_lds(L, S, t, ex, (int)f, p, w.acc);
        }
      }
    }
  }
  acc_lds_flush(L, S, 2, w.acc);
  if (last_block(&w.tickets[1], gridDim.x)) {
    const int turn = e.state->turn;
    const phx_turn_program &g = e.programs[e.turn_program[turn]];
    acc_finalize(w.acc, S, t, ex, e.counts, nullptr, e.slice_means);
    __syncthreads();
    // history entry: times = z_centers / (beta c) (particles/slicing.py:124-126, 303-307)
    phx_wake_hist *h = e.hist;
    __shared__ int sh_head;
    if (threadIdx.x == 0) {
      const int cap = h->capacity, nf = h->n_filled, head = h->head;
      for (int i = 0; i < nf; ++i) {
        const int sl = ring_slot(head, i, cap);
        h->age[sl] = h->age[sl] + g.age_increment;
      }
      const int nh = (head + 1) % cap;
      h->head = nh;
      h->n_filled = nf + 1 < cap ? nf + 1 : cap;
      h->age[nh] = 0.0;
      h->beta[nh] = g.beta;
      sh_head = nh;
    }
    __syncthreads();
    const int nh = sh_head;
    double *tt = h->times + (size_t)nh * S;
    double *mm = h->mom + (size_t)nh * 3 * S;
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
      const double zc = e.z_bins[s] + 0.5 * (e.z_bins[s + 1] - e.z_bins[s]);
      tt[s] = zc / g.beta_c;
      const double cnt = (double)e.counts[s];
      mm[s] = cnt;
      mm[S + s] = cnt * e.slice_means[s];
      mm[2 * S + s] = cnt * e.slice_means[S + s];
    }
  }
}

__global__ void __launch_bounds__(PHX_BLOCK) k_engine_conv(EngineArgs a) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices, a.n_kicks, a.T);
  const phx_turn_program &g = e.programs[e.turn_program[e.state->turn]];
  const WakeKicks &kk = *reinterpret_cast<const WakeKicks *>(g.kick);
  wake_conv_block(e.hist, kk, e.n_slices, w.conv_part, a.T, w.conv_tickets, e.kicks);
  if (blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0) *e.pending = 1;
}

__global__ void k_engine_reset(EngineArgs a, size_t clear_bytes) {
  const phx_engine &e = a.e;
  // clear tickets, partials and accumulators of the workspace
  unsigned long long *p = reinterpret_cast<unsigned long long *>(e.ws);
  const size_t nw = clear_bytes / 8;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nw;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ULL;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *e.turn_counter = 0;
    *e.pending = 0;
    e.hist->n_filled = 0;
    e.hist->head = e.hist->capacity - 1;
    e.state->valid = 0;
  }
}

int check_engine(const phx_engine *e, int *n_kicks, int *T) {
  if (!e) return phx::fail(PHX_E_INVALID, "phx_engine: null engine");
  for (int i = 0; i < 6; ++i)
    if (!e->coords[i]) return phx::fail(PHX_E_INVALID, "phx_engine: coordinate %d is null", i);
  if (e->n < 1) return phx::fail(PHX_E_INVALID, "phx_engine: n < 1");
  if (!e->programs || !e->turn_program || !e->turn_counter || !e->state || !e->ws || !e->pending)
    return phx::fail(PHX_E_INVALID, "phx_engine: null state pointer");
  if (e->n_slices < 0) return phx::fail(PHX_E_INVALID, "phx_engine: n_slices < 0");
  if (e->n_kicks < 0 || e->n_kicks > PHX_MAX_KICKS)
    return phx::fail(PHX_E_LIMIT, "phx_engine: n_kicks %d", e->n_kicks);
  if (e->n_slices > 0 && (e->n_kicks < 1 || e->hist_capacity < 1))
    return phx::fail(PHX_E_INVALID, "phx_engine: wake field needs kicks and history");
  if (e->record_stats && !e->stats) return phx::fail(PHX_E_INVALID, "phx_engine: stats null");
  if (e->n_slices > 0 && (!e->hist || !e->z_bins || !e->kicks || !e->counts || !e->slice_means))
    return phx::fail(PHX_E_INVALID, "phx_engine: wake buffers null");
  *n_kicks = e->n_kicks;
  *T = e->hist_capacity > 0 ? e->hist_capacity : 1;
  return PHX_OK;
}

}  // namespace

extern "C" int64_t phx_engine_ws_bytes(int64_t n, int32_t n_slices, int32_t n_kicks,
                                       int32_t max_turns) {
  if (n < 1 || n_slices < 0 || n_kicks < 0 || max_turns < 1) return -1;
  return (int64_t)engine_ws_bytes(n, n_slices, n_kicks, max_turns);
}

extern "C" int phx_engine_reset(const phx_engine *e, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  EngineArgs a{*e, nk, T};
  const size_t bytes = engine_ws_bytes(e->n, e->n_slices, nk, T);
  k_engine_reset<<<256, 256, 0, phx::as_stream(stream)>>>(a, bytes);
  PHX_CHECK_LAUNCH("phx_engine_reset");
  return PHX_OK;
}

static int enqueue_stage(const EngineArgs &a, int stage, bool has_wake, hipStream_t s) {
  const phx_engine &e = a.e;
  if (stage == 0) {
    const dim3 grid((unsigned)phx::n_chunks(e.n));
    if (has_wake) {
      if (e.record_stats) k_engine_pass<0, true, true><<<grid, PHX_BLOCK, 0, s>>>(a);
      else k_engine_pass<0, false, true><<<grid, PHX_BLOCK, 0, s>>>(a);
    } else {
      if (e.record_stats) k_engine_pass<0, true, false><<<grid, PHX_BLOCK, 0, s>>>(a);
      else k_engine_pass<0, false, false><<<grid, PHX_BLOCK, 0, s>>>(a);
    }
    PHX_CHECK_LAUNCH("phx_engine: pass");
  } else if (stage == 1) {
    const int64_t nc = phx::n_chunks(e.n);
    const unsigned hb = (unsigned)(nc < EN_HIST_BLOCKS ? nc : EN_HIST_BLOCKS);
    k_engine_hist<<<dim3(hb), PHX_BLOCK, acc_lds_bytes(e.n_slices, 2), s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: hist");
  } else {
    const int tiles = (e.n_slices + WK_TILE - 1) / WK_TILE;
    k_engine_conv<<<dim3(tiles, a.n_kicks, a.T), PHX_BLOCK, wake_conv_lds_bytes(e.n_slices),
                    s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: conv");
  }
  return PHX_OK;
}

static int enqueue_turn(const EngineArgs &a, bool has_wake, hipStream_t s) {
  int rc = enqueue_stage(a, 0, has_wake, s);
  if (rc || !has_wake) return rc;
  rc = enqueue_stage(a, 1, has_wake, s);
  if (rc) return rc;
  return enqueue_stage(a, 2, has_wake, s);
}

extern "C" int phx_engine_turns(const phx_engine *e, int32_t n_turns, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  if (n_turns < 0) return phx::fail(PHX_E_INVALID, "phx_engine_turns: n_turns < 0");
  const bool has_wake = e->n_slices > 0 && nk > 0;
  if (has_wake) {
    if (acc_lds_bytes(e->n_slices, 2) > phx_max_lds_bytes() ||
        wake_conv_lds_bytes(e->n_slices) > phx_max_lds_bytes())
      return phx::fail(PHX_E_LIMIT, "phx_engine_turns: n_slices=%d exceeds LDS", e->n_slices);
  }
  EngineArgs a{*e, nk, T};
  for (int t = 0; t < n_turns; ++t) {
    rc = enqueue_turn(a, has_wake, phx::as_stream(stream));
    if (rc) return rc;
  }
  return PHX_OK;
}

extern "C" int phx_engine_stage(const phx_engine *e, int32_t stage, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  const bool has_wake = e->n_slices > 0 && nk > 0;
  if (stage < 0 || stage > 2 || (stage > 0 && !has_wake))
    return phx::fail(PHX_E_INVALID, "phx_engine_stage: bad stage %d", stage);
  EngineArgs a{*e, nk, T};
  return enqueue_stage(a, stage, has_wake, phx::as_stream(stream));
}

extern "C" int phx_engine_flush(const phx_engine *e, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  EngineArgs a{*e, nk, T};
  const dim3 grid((unsigned)phx::n_chunks(e->n));
  if (e->record_stats) k_engine_pass<1, true, false><<<grid, PHX_BLOCK, 0, phx::as_stream(stream)>>>(a);
  else k_engine_pass<1, false, false><<<grid, PHX_BLOCK, 0, phx::as_stream(stream)>>>(a);
  PHX_CHECK_LAUNCH("phx_engine_flush");
  return PHX_OK;
}

struct GraphHandle {
  hipGraphExec_t exec;
};

extern "C" int phx_graph_capture(const phx_engine *e, int32_t n_turns, void **graph) {
  if (!graph) return phx::fail(PHX_E_INVALID, "phx_graph_capture: null output");
  *graph = nullptr;
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  if (n_turns < 1) return phx::fail(PHX_E_INVALID, "phx_graph_capture: n_turns < 1");
  hipStream_t cs;
  PHX_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "phx_graph_capture");
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return phx::fail((int)err, "phx_graph_capture: begin: %s", hipGetErrorString(err));
  }
  rc = phx_engine_turns(e, n_turns, cs);
  err = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return phx::fail((int)err, "phx_graph_capture: end: %s", hipGetErrorString(err));
  hipGraphExec_t ex = nullptr;
  err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess)
    return phx::fail((int)err, "phx_graph_capture: instantiate: %s", hipGetErrorString(err));
  GraphHandle *h = new GraphHandle{ex};
  *graph = h;
  return PHX_OK;
}

extern "C" int phx_graph_launch(void *graph, phx_stream_t stream) {
  if (!graph) return phx::fail(PHX_E_INVALID, "phx_graph_launch: null graph");
  GraphHandle *h = reinterpret_cast<GraphHandle *>(graph);
  PHX_CHECK_HIP(hipGraphLaunch(h->exec, phx::as_stream(stream)), "phx_graph_launch");
  return PHX_OK;
}

extern "C" int phx_graph_destroy(void *graph) {
  if (!graph) return PHX_OK;
  GraphHandle *h = reinterpret_cast<GraphHandle *>(graph);
  hipError_t err = hipGraphExecDestroy(h->exec);
  delete h;
  if (err != hipSuccess) return phx::fail((int)err, "phx_graph_destroy: %s", hipGetErrorString(err));
  return PHX_OK;
}
