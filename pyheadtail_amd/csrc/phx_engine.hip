// Fused one-turn map.  One turn of [local elements..., WakeField] is four
// launches:
//   1. particle pass (persistent grid, the next particle group loading while
//      the current one computes): apply the wake kick left pending by the
//      previous turn (it needs only z and the previous cuts, both unchanged
//      since), record the bunch moments of that end-of-turn state, run the
//      local elements (longitudinal maps, transverse segments), write the six
//      coordinates once (96 B per macroparticle), and store per-block
//      partial sums of the new z moments;
//   2. slice histogram: every block first folds the pass partials in the
//      same fixed order into the cuts (block 0 also writes the statistics row
//      and advances the wake-history ring), then keeps an LDS histogram of
//      counts and fixed-point sums of x and y (24 B per macroparticle) and
//      stores it as its partial;
//   3. merge (slice-parallel): exact totals, slice means, bin edges and the
//      new history entry (times, moments);
//   4. wake convolution over the history (and, with several turns of memory,
//      the in-order sum over turns).
// Without a wake field a turn is the pass and a one-block finalize.
// No stage uses cross-block atomics or fences: kernel boundaries order the
// stages, which keeps every reduction deterministic and avoids the
// serialisation of device-scope atomics.  The kick of turn t is applied in
// pass 1 of turn t+1, which gives the arithmetic of WakeField.track at the
// end of turn t (impedances/wakes.py:121-145).
#include "phx_engine.h"
#include "phx_host.h"

using namespace phx_en;

namespace {
template <int MODE, bool REC, bool WAKE, int SHAPE, int KS, int DS>
__global__ void __launch_bounds__(PHX_BLOCK, PHX_PASS_MIN_BLOCKS)
    k_engine_pass(EngineArgs a, double *__restrict__ x, double *__restrict__ xp,
                  double *__restrict__ y, double *__restrict__ yp, double *__restrict__ z,
                  double *__restrict__ dp, const phx_turn_program *__restrict__ programs,
                  const double *__restrict__ kicks, int stage) {
  pass_body<MODE, REC, WAKE, SHAPE, KS, DS>(a, x, xp, y, yp, z, dp, programs, kicks, blockIdx.x,
                                            gridDim.x, stage);
}

// PHX_SCPASS_WAVES > 0: ask the compiler for that many waves per SIMD (it
// then keeps the registers within the budget, spilling if it must).
#ifndef PHX_SCPASS_WAVES
#define PHX_SCPASS_WAVES 0
#endif
#if PHX_SCPASS_WAVES > 0
#define PHX_SCPASS_ATTR __attribute__((amdgpu_waves_per_eu(PHX_SCPASS_WAVES)))
#else
#define PHX_SCPASS_ATTR
#endif
template <int SHAPE, int DS>
__global__ void __launch_bounds__(PHX_BLOCK) PHX_SCPASS_ATTR
    k_engine_scpass(EngineArgs a, double *__restrict__ x, double *__restrict__ xp,
                    double *__restrict__ y, double *__restrict__ yp, double *__restrict__ z,
                    double *__restrict__ dp, const phx_turn_program *__restrict__ programs,
                    int stage) {
  sc_pass_body<SHAPE, DS>(a, x, xp, y, yp, z, dp, programs, blockIdx.x, gridDim.x, stage);
}

// One block: fold the chunk partials, then the per-turn bookkeeping (for
// maps without a wake field, and for the flush).  ZBL: the partials have
// the z-block layout (the passes of space-charge engines).
template <int MODE, bool REC, bool ZBL = false>
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_fin(EngineArgs a) {
  using L = PassLayout<REC, ZBL>;
  const phx_engine &e = a.e;
  __shared__ double res[L::NV];
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int turn = *e.turn_counter;
  fold_partials<L::MINM, L::MAXM>(w.partials, a.grid, 0, L::NV, res);
  if (REC && threadIdx.x < EN_STAT_NV)
    e.stats[(size_t)turn * EN_STAT_NV + threadIdx.x] = res[L::SOFF + threadIdx.x];
  if (threadIdx.x < 6) w.p0[threadIdx.x] = e.coords[threadIdx.x][0];  // the next pass's shifts
  if (threadIdx.x == 0) {
    if (MODE == 0) *e.turn_counter = turn + 1;
    else *e.pending = 0;
  }
}

// Particle 0 as it is now, the shifts of the next pass.  The kernels that
// follow a pass (finalize, histogram) refresh it; this one is for the first
// pass of a launch sequence, after which the coordinates may have changed.
__global__ void k_engine_snapshot(EngineArgs a) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  if (threadIdx.x < 6) w.p0[threadIdx.x] = e.coords[threadIdx.x][0];
}

// The pending wake kick alone (a flush without statistics): one pass over
// the coordinates the kick set reads, writing the momenta it changes
// (16 B of z and 32 B of xp, yp per macroparticle for dipolar kicks).
template <int KS>
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_kick(EngineArgs a) {
  const phx_engine &e = a.e;
  if (!*e.pending) return;
  const int turn = *e.turn_counter;
  const phx_turn_program &g = e.programs[e.turn_program[turn > 0 ? turn - 1 : 0]];
  const int nk = g.n_kicks < a.n_kicks ? g.n_kicks : a.n_kicks;
  const double tail = e.state->tail, width = e.state->width;
  const double inv_width = 1.0 / width;
  const int S = e.n_slices;
  double *const *c = e.coords;
  for (int64_t j = blockIdx.x * (int64_t)PHX_BLOCK + threadIdx.x; j < e.n;
       j += (int64_t)gridDim.x * PHX_BLOCK) {
    P6 p;
    p.z = c[4][j];
    p.xp = c[1][j];
    p.yp = c[3][j];
    p.x = KS == PHX_KICKS_DXDY ? 0.0 : c[0][j];
    p.y = KS == PHX_KICKS_DXDY ? 0.0 : c[2][j];
    p.dp = KS == PHX_KICKS_GENERIC ? c[5][j] : 0.0;
    const double f = slice_floor_r(p.z, tail, width, inv_width);
    if (f >= 0.0 && f < (double)S) {
      apply_kick_set<KS>(p, (int32_t)f, S, g.kick, nk, e.kicks);
      c[1][j] = p.xp;
      c[3][j] = p.yp;
      if (KS == PHX_KICKS_GENERIC) c[5][j] = p.dp;
    }
  }
}

template <bool REC>
__global__ void __launch_bounds__(EN_HIST_THREADS) k_engine_hist(EngineArgs a) {
  hist_body<REC, true>(a, blockIdx.x, gridDim.x);
}

template <int SL>
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_merge(EngineArgs a) {
  merge_body<SL>(a, blockIdx.x);
}

__global__ void __launch_bounds__(WK_THREADS) k_engine_conv(EngineArgs a) {
  conv_body(a, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <int KM>
__global__ void __launch_bounds__(rw_threads<KM>()) k_engine_conv_rw(EngineArgs a) {
  conv_rw_body<KM>(a, blockIdx.x);
}

__global__ void k_engine_conv_reduce(EngineArgs a) { conv_reduce_body(a, blockIdx.x); }

__global__ void __launch_bounds__(EN_HIST_THREADS) k_engine_schist(EngineArgs a, int stage) {
  sc_hist_body(a, stage, blockIdx.x, gridDim.x);
}

template <int SL>
__global__ void __launch_bounds__(PHX_BLOCK) k_engine_smerge(EngineArgs a, int stage) {
  sc_merge_body<SL>(a, stage, blockIdx.x);
}

// The space-charge workspace at reset: zeroed, then the Faddeeva
// coefficients.
__global__ void k_engine_sc_init(EngineArgs a, WofzCoeffs c, size_t clear_bytes) {
  const phx_engine &e = a.e;
  unsigned long long *p = reinterpret_cast<unsigned long long *>(e.sc_ws);
  const size_t nw = clear_bytes / 8;
  const size_t first = sc_r256(sizeof(WofzCoeffs)) / 8;
  for (size_t i = first + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nw;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ULL;
  if (blockIdx.x == 0 && threadIdx.x == 0) *reinterpret_cast<WofzCoeffs *>(e.sc_ws) = c;
}

__global__ void k_engine_reset(EngineArgs a, size_t clear_bytes) {
  const phx_engine &e = a.e;
  unsigned long long *p = reinterpret_cast<unsigned long long *>(e.ws);
  const size_t nw = clear_bytes / 8;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nw;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ULL;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *e.turn_counter = 0;
    *e.pending = 0;
    if (e.hist) {
      e.hist->n_filled = 0;
      e.hist->head = e.hist->capacity - 1;
    }
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
  {
    const int base = e->program_shape & ~PHX_SHAPE_STATIONARY;
    if (base < 0 || base >= PHX_SHAPE_COUNT ||
        (base != e->program_shape && base != PHX_SHAPE_RF1_T))
      return phx::fail(PHX_E_INVALID, "phx_engine: program_shape %d", e->program_shape);
  }
  if (e->kick_shape < 0 || e->kick_shape >= PHX_KICKS_COUNT)
    return phx::fail(PHX_E_INVALID, "phx_engine: kick_shape %d", e->kick_shape);
  if (e->pass_grid < 0 || e->pass_grid > PHX_MAX_GRID)
    return phx::fail(PHX_E_INVALID, "phx_engine: pass_grid %d not in 0..%d", e->pass_grid,
                     PHX_MAX_GRID);
  if (e->conv_kinds < 0 || e->conv_kinds > (PHX_CONV_GENERAL | PHX_CONV_RW_DOT | PHX_CONV_RW_MFMA))
    return phx::fail(PHX_E_INVALID, "phx_engine: conv_kinds %d", e->conv_kinds);
  if (e->kick_shape == PHX_KICKS_DXDY && e->n_kicks != 2)
    return phx::fail(PHX_E_INVALID, "phx_engine: kick_shape DXDY needs 2 kicks, not %d",
                     e->n_kicks);
  if (e->kick_shape == PHX_KICKS_PLATES && e->n_kicks != 4)
    return phx::fail(PHX_E_INVALID, "phx_engine: kick_shape PLATES needs 4 kicks, not %d",
                     e->n_kicks);
  {
    const int ds = e->det_shape & ~PHX_DETS_SMOOTH;
    if (ds < 0 || ds >= PHX_DETS_COUNT || (ds < PHX_DETS_C1 && ds != e->det_shape))
      return phx::fail(PHX_E_INVALID, "phx_engine: det_shape %d", e->det_shape);
  }
  // the particle pass addresses a coordinate by a 32-bit byte offset
  if (e->n > PHX_ENGINE_MAX_N)
    return phx::fail(PHX_E_LIMIT, "phx_engine: %lld macroparticles, the fused engine takes at "
                     "most %lld per bunch", (long long)e->n, (long long)PHX_ENGINE_MAX_N);
  if (e->n_slices > 0 && (!e->hist || !e->z_bins || !e->kicks || !e->counts || !e->slice_means))
    return phx::fail(PHX_E_INVALID, "phx_engine: wake buffers null");
  if (e->sc_stages < 0 || e->sc_stages > PHX_MAX_SC_STAGES)
    return phx::fail(PHX_E_LIMIT, "phx_engine: sc_stages %d not in 0..%d", e->sc_stages,
                     PHX_MAX_SC_STAGES);
  if (e->sc_stages > 0) {
    if (!e->sc_ws || e->sc_slices < 1)
      return phx::fail(PHX_E_INVALID, "phx_engine: space-charge stages need sc_ws and sc_slices");
    for (int k = 1; k <= e->sc_stages; ++k)
      if (e->sc_stage_shape[k] < 0 || e->sc_stage_shape[k] >= PHX_SHAPE_COUNT)
        return phx::fail(PHX_E_INVALID, "phx_engine: sc_stage_shape[%d] = %d", k,
                         e->sc_stage_shape[k]);
  }
  *n_kicks = e->n_kicks;
  *T = e->hist_capacity > 0 ? e->hist_capacity : 1;
  return PHX_OK;
}

template <bool REC, bool WAKE>
size_t pass_lds_bytes() {
  constexpr int NV = PassLayout<REC, WAKE>::NV;
  return BR_LDS(NV) * sizeof(double);
}

template <int MODE, bool R, bool W, int SH, int KS, int DS>
void launch_pass_kernel(const EngineArgs &a, int stage, hipStream_t s) {
  const phx_engine &e = a.e;
  double *const *c = e.coords;
  k_engine_pass<MODE, R, W, SH, KS, DS>
      <<<dim3((unsigned)a.grid), PHX_BLOCK, pass_lds_bytes<R, W>(), s>>>(
          a, c[0], c[1], c[2], c[3], c[4], c[5], e.programs, e.kicks, stage);
}

// Blocks of this pass that stay resident on one CU (queried once).
template <int MODE, bool R, bool W, int SH, int KS, int DS>
int pass_blocks_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &nb, k_engine_pass<MODE, R, W, SH, KS, DS>, PHX_BLOCK,
            pass_lds_bytes<R, W>()) != hipSuccess ||
        nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}

struct PassKernel {
  void (*launch)(const EngineArgs &, int, hipStream_t);
  int (*blocks_per_cu)();
};

template <int MODE, bool R, bool W, int SH, int KS, int DS>
PassKernel pass_kernel() {
  return PassKernel{launch_pass_kernel<MODE, R, W, SH, KS, DS>,
                    pass_blocks_per_cu<MODE, R, W, SH, KS, DS>};
}

// The detuner set compiled in straight-line code: turns of the specialised
// shapes without statistics records (the statistics runs take the generic
// detuner loop, the same arithmetic).
template <int MODE, bool R, bool W, int SH, int KS>
PassKernel pick_dets(int ds) {
  if constexpr (MODE == 0 && !R && SH != PHX_SHAPE_GENERIC) {
    switch (ds) {
      case PHX_DETS_NONE: return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_NONE>();
      case PHX_DETS_C1: return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_C1>();
      case PHX_DETS_C1_AMP: return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_C1_AMP>();
      case PHX_DETS_C1 | PHX_DETS_SMOOTH:
        return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_C1 | PHX_DETS_SMOOTH>();
      case PHX_DETS_C1_AMP | PHX_DETS_SMOOTH:
        return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_C1_AMP | PHX_DETS_SMOOTH>();
      default: break;
    }
  }
  return pass_kernel<MODE, R, W, SH, KS, PHX_DETS_GENERIC>();
}

// The pass compiled for (program shape, kick set, detuner set); a flush
// applies no program and runs the generic one.
template <int MODE, bool R, bool W, int SH>
PassKernel pick_kicks(int ks, int ds) {
  if constexpr (W && MODE == 0) {
    if (ks == PHX_KICKS_DXDY) return pick_dets<MODE, R, W, SH, PHX_KICKS_DXDY>(ds);
    if (ks == PHX_KICKS_PLATES) return pick_dets<MODE, R, W, SH, PHX_KICKS_PLATES>(ds);
  }
  return pick_dets<MODE, R, W, SH, PHX_KICKS_GENERIC>(ds);
}

template <int MODE, bool R, bool W>
PassKernel pick_shape(int shape, int ks, int ds) {
  if constexpr (MODE == 0) {
    switch (shape) {
      case PHX_SHAPE_RF1_T | PHX_SHAPE_STATIONARY:
        // compiled for the wake pass with the straight-line kick sets;
        // otherwise the general RF code, which gives the same bits
        if constexpr (W && !R) {
          constexpr int ST = PHX_SHAPE_RF1_T | PHX_SHAPE_STATIONARY;
          if (ks == PHX_KICKS_DXDY) return pick_dets<MODE, R, W, ST, PHX_KICKS_DXDY>(ds);
          if (ks == PHX_KICKS_PLATES) return pick_dets<MODE, R, W, ST, PHX_KICKS_PLATES>(ds);
        }
        return pick_kicks<MODE, R, W, PHX_SHAPE_RF1_T>(ks, ds);
      case PHX_SHAPE_RF1_T: return pick_kicks<MODE, R, W, PHX_SHAPE_RF1_T>(ks, ds);
      case PHX_SHAPE_RF2_T: return pick_kicks<MODE, R, W, PHX_SHAPE_RF2_T>(ks, ds);
      case PHX_SHAPE_T_RF1: return pick_kicks<MODE, R, W, PHX_SHAPE_T_RF1>(ks, ds);
      case PHX_SHAPE_T_RF2: return pick_kicks<MODE, R, W, PHX_SHAPE_T_RF2>(ks, ds);
      case PHX_SHAPE_LIN_T: return pick_kicks<MODE, R, W, PHX_SHAPE_LIN_T>(ks, ds);
      case PHX_SHAPE_T_LIN: return pick_kicks<MODE, R, W, PHX_SHAPE_T_LIN>(ks, ds);
      case PHX_SHAPE_T: return pick_kicks<MODE, R, W, PHX_SHAPE_T>(ks, ds);
      default: break;
    }
  }
  return pick_kicks<MODE, R, W, PHX_SHAPE_GENERIC>(ks, ds);
}

// MODE 0 with a wake field: the histogram kernel folds the pass's partials;
// otherwise a one-block finalize does.
template <int MODE>
PassKernel pick_pass(const phx_engine &e, bool wake_pass) {
  const bool rec = e.record_stats != 0;
  const int shape = e.program_shape, ks = e.kick_shape, ds = e.det_shape;
  if (MODE == 0 && wake_pass)
    return rec ? pick_shape<MODE, true, true>(shape, ks, ds)
               : pick_shape<MODE, false, true>(shape, ks, ds);
  return rec ? pick_shape<MODE, true, false>(shape, ks, ds)
             : pick_shape<MODE, false, false>(shape, ks, ds);
}

// The pass grid: the caller's, or as many blocks as stay resident.
int pass_grid(const phx_engine &e, const PassKernel &k) {
  int64_t g = e.pass_grid > 0 ? (int64_t)e.pass_grid
                              : (int64_t)k.blocks_per_cu() * (int64_t)phx_num_cus();
  if (g > PHX_MAX_GRID) g = PHX_MAX_GRID;
  if (g < 1) g = 1;
  return phx_grid_for(e.n, g);
}

// The space-charge stage pass compiled for a stage's element order and the
// detuner set (the orders the served rings have; others: the generic pass).
struct ScPassKernel {
  void (*kernel)(EngineArgs, double *, double *, double *, double *, double *, double *,
                 const phx_turn_program *, int);
  int (*blocks_per_cu)();
};

template <int SH, int DS>
int scpass_blocks_per_cu() {
  static int cached = 0;
  if (cached == 0) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_engine_scpass<SH, DS>, PHX_BLOCK,
                                                     BR_LDS(ZB_NV) * sizeof(double)) !=
            hipSuccess ||
        nb < 1)
      nb = 1;
    cached = nb;
  }
  return cached;
}

template <int SH, int DS>
ScPassKernel scpass_kernel() {
  return ScPassKernel{k_engine_scpass<SH, DS>, scpass_blocks_per_cu<SH, DS>};
}

template <int SH>
ScPassKernel scpass_dets(int ds) {
  switch (ds) {
    case PHX_DETS_NONE: return scpass_kernel<SH, PHX_DETS_NONE>();
    case PHX_DETS_C1: return scpass_kernel<SH, PHX_DETS_C1>();
    case PHX_DETS_C1_AMP: return scpass_kernel<SH, PHX_DETS_C1_AMP>();
    case PHX_DETS_C1 | PHX_DETS_SMOOTH: return scpass_kernel<SH, PHX_DETS_C1 | PHX_DETS_SMOOTH>();
    case PHX_DETS_C1_AMP | PHX_DETS_SMOOTH:
      return scpass_kernel<SH, PHX_DETS_C1_AMP | PHX_DETS_SMOOTH>();
    default: return scpass_kernel<SH, PHX_DETS_GENERIC>();
  }
}

ScPassKernel pick_scpass(int shape, int ds) {
  switch (shape) {
    case PHX_SHAPE_T: return scpass_dets<PHX_SHAPE_T>(ds);
    case PHX_SHAPE_LIN_T: return scpass_dets<PHX_SHAPE_LIN_T>(ds);
    case PHX_SHAPE_T_LIN: return scpass_dets<PHX_SHAPE_T_LIN>(ds);
    case PHX_SHAPE_NONE: return scpass_kernel<PHX_SHAPE_NONE, PHX_DETS_GENERIC>();
    default: return scpass_kernel<PHX_SHAPE_GENERIC, PHX_DETS_GENERIC>();
  }
}

// Resident blocks per CU of the least occupying stage pass of an engine.
int scpass_blocks_per_cu(const phx_engine &e) {
  int nb = 1 << 20;
  for (int k = 1; k <= e.sc_stages; ++k) {
    const int b = pick_scpass(e.sc_stage_shape[k], e.det_shape).blocks_per_cu();
    nb = b < nb ? b : nb;
  }
  return nb;
}

EngineArgs make_args(const phx_engine *e, int nk, int T) {
  const bool has_wake = e->n_slices > 0 && nk > 0;
  EngineArgs a{*e, nk, T, 1, 1};
  if (e->sc_stages > 0 && e->pass_grid == 0) {
    // every pass of a turn shares the grid (the partials' layout); space-
    // charge engines take the one that fills the device with their stage
    // passes, which carry most of the work
    int64_t g = (int64_t)scpass_blocks_per_cu(*e) * (int64_t)phx_num_cus();
    if (g > PHX_MAX_GRID) g = PHX_MAX_GRID;
    a.grid = phx_grid_for(e->n, g);
  } else {
    a.grid = pass_grid(*e, pick_pass<0>(*e, has_wake || e->sc_stages > 0));
  }
  if (e->n_slices > 0) a.gmax = wake_conv_gmax(e->n_slices, phx_max_lds_bytes());
  return a;
}

// The space-charge stages of a turn (sc_stages = m): for k = 0 .. m, the
// pass of stage k (stage 0: the pending wake kick, statistics and the
// turn's first elements; stage k > 0: the space-charge kick ending stage
// k-1, then stage k's elements), each but the last followed by the
// space-charge slicing of its result (histogram and merge).  The z block of
// the last pass feeds the wake field's cuts, or the finalize without one.
int launch_sc_stages(const EngineArgs &a, bool has_wake, hipStream_t s) {
  const phx_engine &e = a.e;
  const int m = e.sc_stages;
  const int S = e.sc_slices;
  const unsigned hb = (unsigned)hist_blocks(e.n);
  const size_t lds = acc_lds_bytes(S, SC_TERMS);
  for (int k = 0; k <= m; ++k) {
    if (k == 0) {
      pick_pass<0>(e, true).launch(a, 0, s);
    } else {
      double *const *c = e.coords;
      const ScPassKernel kk = pick_scpass(e.sc_stage_shape[k], e.det_shape);
      hipLaunchKernelGGL(kk.kernel, dim3((unsigned)a.grid), dim3(PHX_BLOCK),
                         BR_LDS(ZB_NV) * sizeof(double), s, a, c[0], c[1], c[2], c[3], c[4],
                         c[5], e.programs, k);
    }
    PHX_CHECK_LAUNCH("phx_engine: space-charge stage pass");
    if (k == m) break;
    k_engine_schist<<<dim3(hb), EN_HIST_THREADS, lds, s>>>(a, k);
    PHX_CHECK_LAUNCH("phx_engine: space-charge histogram");
    if (merge_sl(S) == 4)
      k_engine_smerge<4><<<dim3((S + 3) / 4), PHX_BLOCK, 0, s>>>(a, k);
    else
      k_engine_smerge<8><<<dim3((S + 7) / 8), PHX_BLOCK, 0, s>>>(a, k);
    PHX_CHECK_LAUNCH("phx_engine: space-charge merge");
  }
  if (has_wake) return PHX_OK;
  if (e.record_stats) k_engine_fin<0, true, true><<<1, PHX_BLOCK, 0, s>>>(a);
  else k_engine_fin<0, false, true><<<1, PHX_BLOCK, 0, s>>>(a);
  PHX_CHECK_LAUNCH("phx_engine: finalize");
  return PHX_OK;
}

template <int MODE>
int launch_pass(const EngineArgs &args, bool has_wake, hipStream_t s) {
  if (MODE == 0 && args.e.sc_stages > 0) return launch_sc_stages(args, has_wake, s);
  const bool wake_pass = MODE == 0 && has_wake;
  const PassKernel k = pick_pass<MODE>(args.e, wake_pass);
  EngineArgs a = args;
  if (MODE != 0) a.grid = pass_grid(a.e, k);
  k.launch(a, 0, s);
  PHX_CHECK_LAUNCH("phx_engine: pass");
  if (wake_pass) return PHX_OK;
  if (a.e.record_stats) k_engine_fin<MODE, true><<<1, PHX_BLOCK, 0, s>>>(a);
  else k_engine_fin<MODE, false><<<1, PHX_BLOCK, 0, s>>>(a);
  PHX_CHECK_LAUNCH("phx_engine: finalize");
  return PHX_OK;
}

int enqueue_stage(const EngineArgs &a, int stage, bool has_wake, hipStream_t s) {
  const phx_engine &e = a.e;
  if (stage == 0) return launch_pass<0>(a, has_wake, s);
  if (stage == 1) {
    if (e.record_stats)
      k_engine_hist<true><<<dim3(hist_blocks(e.n)), EN_HIST_THREADS, acc_lds_bytes(e.n_slices, 2), s>>>(a);
    else
      k_engine_hist<false><<<dim3(hist_blocks(e.n)), EN_HIST_THREADS, acc_lds_bytes(e.n_slices, 2), s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: hist");
    // one more block describes the resistive-wall kicks for the convolution
    const int desc = has_wake && engine_has_rw(e) ? 1 : 0;
    if (merge_sl(e.n_slices) == 4)
      k_engine_merge<4><<<dim3((e.n_slices + 3) / 4 + desc), PHX_BLOCK, 0, s>>>(a);
    else
      k_engine_merge<8><<<dim3((e.n_slices + 7) / 8 + desc), PHX_BLOCK, 0, s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: merge");
    return PHX_OK;
  }
  if (e.conv_kinds == 0 || (e.conv_kinds & PHX_CONV_GENERAL)) {
    const int tiles = (e.n_slices + WK_TILE - 1) / WK_TILE;
    k_engine_conv<<<dim3(tiles, a.n_kicks, a.T), WK_THREADS,
                    wake_conv_lds_bytes(e.n_slices, a.gmax), s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: conv");
  }
  if (e.conv_kinds == 0 || (e.conv_kinds & PHX_CONV_RW_DOT)) {
    k_engine_conv_rw<0><<<dim3((e.n_slices + RW_TILE - 1) / RW_TILE * conv_rows(e.n_slices, a.T)), rw_threads<0>(),
                          wake_rw_lds_bytes(e.n_slices), s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: conv resistive wall");
  }
  if (e.conv_kinds == 0 || (e.conv_kinds & PHX_CONV_RW_MFMA)) {
    k_engine_conv_rw<1><<<dim3((e.n_slices + RW_TILE - 1) / RW_TILE * conv_rows(e.n_slices, a.T)), rw_threads<1>(),
                          wake_rw_lds_bytes(e.n_slices), s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: conv resistive wall (matrix cores)");
  }
  if (conv_reduces(e.n_slices, a.T, engine_has_rw(e))) {
    const int tot = a.n_kicks * e.n_slices;
    k_engine_conv_reduce<<<dim3((tot * RED_LANES + 255) / 256), 256, 0, s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine: conv reduce");
  }
  return PHX_OK;
}

int enqueue_snapshot(const EngineArgs &a, hipStream_t s) {
  k_engine_snapshot<<<1, 64, 0, s>>>(a);
  PHX_CHECK_LAUNCH("phx_engine: snapshot");
  return PHX_OK;
}

// The pending kick of the last turn: without statistics, the kick alone
// and the one-block bookkeeping; with them, a snapshot and the flush pass
// (kick, then the final statistics row).
int enqueue_flush(const EngineArgs &a, hipStream_t s) {
  const phx_engine &e = a.e;
  if (e.record_stats || e.n_slices == 0) {
    int rc = enqueue_snapshot(a, s);
    if (rc) return rc;
    return launch_pass<1>(a, false, s);
  }
  const int64_t ng = (e.n + PHX_BLOCK - 1) / PHX_BLOCK;
  const unsigned grid = (unsigned)(ng < PHX_MAX_GRID * 4 ? ng : PHX_MAX_GRID * 4);
  if (e.kick_shape == PHX_KICKS_DXDY)
    k_engine_kick<PHX_KICKS_DXDY><<<grid, PHX_BLOCK, 0, s>>>(a);
  else if (e.kick_shape == PHX_KICKS_PLATES)
    k_engine_kick<PHX_KICKS_PLATES><<<grid, PHX_BLOCK, 0, s>>>(a);
  else
    k_engine_kick<PHX_KICKS_GENERIC><<<grid, PHX_BLOCK, 0, s>>>(a);
  PHX_CHECK_LAUNCH("phx_engine: flush kick");
  k_engine_fin<1, false><<<1, PHX_BLOCK, 0, s>>>(a);
  PHX_CHECK_LAUNCH("phx_engine: flush finalize");
  return PHX_OK;
}

int enqueue_turn(const EngineArgs &a, bool has_wake, hipStream_t s) {
  int rc = enqueue_stage(a, 0, has_wake, s);
  if (rc || !has_wake) return rc;
  rc = enqueue_stage(a, 1, has_wake, s);
  if (rc) return rc;
  return enqueue_stage(a, 2, has_wake, s);
}

int check_lds(const phx_engine *e) {
  if (e->sc_stages > 0 && acc_lds_bytes(e->sc_slices, SC_TERMS) > phx_max_lds_bytes())
    return phx::fail(PHX_E_LIMIT, "phx_engine: sc_slices=%d exceeds LDS", e->sc_slices);
  if (e->n_slices > 0 && (acc_lds_bytes(e->n_slices, 2) > phx_max_lds_bytes() ||
                          wake_conv_lds_bytes(e->n_slices, 1) + WK_STATIC_LDS > phx_max_lds_bytes()))
    return phx::fail(PHX_E_LIMIT, "phx_engine: n_slices=%d exceeds LDS", e->n_slices);
  const bool rw = e->conv_kinds == 0 || (e->conv_kinds & (PHX_CONV_RW_DOT | PHX_CONV_RW_MFMA));
  if (e->n_slices > 0 && rw && wake_rw_lds_bytes(e->n_slices) + RW_STATIC_LDS > phx_max_lds_bytes())
    return phx::fail(PHX_E_LIMIT,
                     "phx_engine: n_slices=%d exceeds LDS of the resistive-wall convolution",
                     e->n_slices);
  return PHX_OK;
}


}  // namespace

int phx_en::engine_args_for(const phx_engine *e, EngineArgs *out) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  rc = check_lds(e);
  if (rc) return rc;
  *out = make_args(e, nk, T);
  return PHX_OK;
}

extern "C" int64_t phx_engine_ws_bytes(int64_t n, int32_t n_slices, int32_t n_kicks,
                                       int32_t max_turns) {
  if (n < 1 || n_slices < 0 || n_kicks < 0 || max_turns < 1) return -1;
  return (int64_t)engine_ws_bytes(n, n_slices, n_kicks, max_turns);
}

extern "C" int64_t phx_engine_sc_ws_bytes(int64_t n, int32_t sc_slices) {
  if (n < 1 || sc_slices < 1) return -1;
  return (int64_t)sc_ws_bytes(n, sc_slices);
}

extern "C" int phx_engine_reset(const phx_engine *e, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  EngineArgs a = make_args(e, nk, T);
  const size_t bytes = engine_ws_bytes(e->n, e->n_slices, nk, T);
  k_engine_reset<<<256, 256, 0, phx::as_stream(stream)>>>(a, bytes);
  PHX_CHECK_LAUNCH("phx_engine_reset");
  if (e->sc_stages > 0) {
    k_engine_sc_init<<<256, 256, 0, phx::as_stream(stream)>>>(
        a, phx_wofz_host_coeffs(), sc_ws_bytes(e->n, e->sc_slices));
    PHX_CHECK_LAUNCH("phx_engine_reset: space charge");
  }
  return PHX_OK;
}

// n_turns turns, after a snapshot of particle 0 if snapshot is set.
static int engine_turns(const phx_engine *e, int32_t n_turns, hipStream_t s, bool snapshot) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  if (n_turns < 0) return phx::fail(PHX_E_INVALID, "phx_engine_turns: n_turns < 0");
  rc = check_lds(e);
  if (rc) return rc;
  const bool has_wake = e->n_slices > 0 && nk > 0;
  EngineArgs a = make_args(e, nk, T);
  if (snapshot && n_turns > 0) {
    rc = enqueue_snapshot(a, s);
    if (rc) return rc;
  }
  for (int t = 0; t < n_turns; ++t) {
    rc = enqueue_turn(a, has_wake, s);
    if (rc) return rc;
  }
  return PHX_OK;
}

extern "C" int phx_engine_turns(const phx_engine *e, int32_t n_turns, phx_stream_t stream) {
  return engine_turns(e, n_turns, phx::as_stream(stream), true);
}

extern "C" int phx_engine_stage(const phx_engine *e, int32_t stage, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  rc = check_lds(e);
  if (rc) return rc;
  const bool has_wake = e->n_slices > 0 && nk > 0;
  if (stage < 0 || stage > 2 || (stage > 0 && !has_wake))
    return phx::fail(PHX_E_INVALID, "phx_engine_stage: bad stage %d", stage);
  EngineArgs a = make_args(e, nk, T);
  return enqueue_stage(a, stage, has_wake, phx::as_stream(stream));
}

extern "C" int phx_engine_flush(const phx_engine *e, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  return enqueue_flush(make_args(e, nk, T), phx::as_stream(stream));
}

extern "C" int phx_engine_snapshot(const phx_engine *e, phx_stream_t stream) {
  int nk, T;
  int rc = check_engine(e, &nk, &T);
  if (rc) return rc;
  return enqueue_snapshot(make_args(e, nk, T), phx::as_stream(stream));
}

// ---------------------------------------------------------------------------
// Allocation-free engine calls (phx_engine_turns_noalloc).  Their workspace
// is this file-scope device array: each device holds its own copy from
// module load until the process ends.  Host code never hands the array's
// name to a kernel (on the host it is only a stand-in symbol); its device
// address comes from hipGetSymbolAddress, once per device at phx_init with
// that device current, and is kept per device in g_scratch.
// ---------------------------------------------------------------------------
#define PHX_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[PHX_NOALLOC_SCRATCH_BYTES];

namespace {
void *g_scratch[PHX_MAX_DEVICES];

// Zeroes the FP64 fallback accumulators of a workspace layout, the only
// part of the workspace a turn reads before writing it (the merge leaves
// them zero again; another engine's call may have left the scratch in any
// state).
__global__ void k_engine_clear_fallback(EngineArgs a) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int nf = 2 * (e.n_slices > 0 ? e.n_slices : 1);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x)
    w.acc.fallback[i] = 0.0;
}
}  // namespace

int phx_engine_scratch_init(int device) {
  if (device < 0 || device >= PHX_MAX_DEVICES)
    return phx::fail(PHX_E_LIMIT, "phx_init: device %d beyond %d", device, PHX_MAX_DEVICES);
  if (g_scratch[device]) return PHX_OK;
  void *p = nullptr;
  const hipError_t err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (err != hipSuccess || !p)
    return phx::fail(err != hipSuccess ? (int)err : PHX_E_STATE,
                     "phx_init: address of the engine scratch on device %d: %s", device,
                     hipGetErrorString(err));
  g_scratch[device] = p;
  return PHX_OK;
}

extern "C" int phx_engine_turns_noalloc(const phx_engine *e, int32_t n_turns, int32_t flags,
                                        phx_stream_t stream) {
  if (!e) return phx::fail(PHX_E_INVALID, "phx_engine_turns_noalloc: null engine");
  if (flags & ~(PHX_NOALLOC_RESET | PHX_NOALLOC_FLUSH))
    return phx::fail(PHX_E_INVALID, "phx_engine_turns_noalloc: flags %d", flags);
  if (n_turns < 0) return phx::fail(PHX_E_INVALID, "phx_engine_turns_noalloc: n_turns < 0");
  int dev = -1;
  PHX_CHECK_HIP(hipGetDevice(&dev), "phx_engine_turns_noalloc");
  if (dev < 0 || dev >= PHX_MAX_DEVICES || !g_scratch[dev])
    return phx::fail(PHX_E_NODEVICE, "phx_engine_turns_noalloc: phx_init(%d) not called", dev);
  if (e->sc_stages != 0)
    return phx::fail(PHX_E_INVALID,
                     "phx_engine_turns_noalloc: space-charge stages need their own workspace "
                     "(phx_engine_turns)");
  phx_engine c = *e;
  c.ws = g_scratch[dev];
  int nk, T;
  int rc = check_engine(&c, &nk, &T);
  if (rc) return rc;
  rc = check_lds(&c);
  if (rc) return rc;
  const size_t need = engine_ws_bytes(c.n, c.n_slices, nk, T);
  if (need > (size_t)PHX_NOALLOC_SCRATCH_BYTES)
    return phx::fail(PHX_E_LIMIT,
                     "phx_engine_turns_noalloc: %lld macroparticles, %d slices, %d kicks and "
                     "%d turns of memory need %zu bytes of scratch; the limit is %d",
                     (long long)c.n, c.n_slices, nk, T, need, PHX_NOALLOC_SCRATCH_BYTES);
  hipStream_t s = phx::as_stream(stream);
  EngineArgs a = make_args(&c, nk, T);
  if (flags & PHX_NOALLOC_RESET) {
    k_engine_reset<<<256, 256, 0, s>>>(a, need);
    PHX_CHECK_LAUNCH("phx_engine_turns_noalloc: reset");
  } else {
    k_engine_clear_fallback<<<8, 256, 0, s>>>(a);
    PHX_CHECK_LAUNCH("phx_engine_turns_noalloc: clear");
  }
  rc = engine_turns(&c, n_turns, s, true);
  if (rc) return rc;
  if (flags & PHX_NOALLOC_FLUSH) return enqueue_flush(a, s);
  return PHX_OK;
}

using phx::GraphHandle;

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
  rc = engine_turns(e, n_turns, cs, false);  // replays continue from the device state
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
