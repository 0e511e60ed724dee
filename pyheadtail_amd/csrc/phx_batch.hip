// Batches of fused engines: many independent bunches (any macroparticle
// count, slice count, wake memory and kick set) advanced together, each
// stage of a turn one launch for all of them.  A launch's blocks are the
// concatenation of every engine's blocks of that stage; block b finds its
// engine by bisection over the stage's block offsets and runs the same
// device body as the single-engine kernel with its local block index, so
// each engine's arithmetic, reduction order included, is that of its run
// alone.  The cuts are folded once per engine by a one-block-per-engine
// stage (the single-engine histogram folds them in every block).
//   stages of a turn: pass | cuts | histogram | merge | convolution
//   (general, resistive wall, resistive wall on the matrix cores) | sum over
//   turns of memory; the empty ones are not launched.
#include <vector>

#include "phx_engine.h"
#include "phx_host.h"

using namespace phx_en;

namespace {

enum { BS_PASS = 0, BS_CUTS, BS_HIST, BS_MERGE, BS_CONV, BS_RW, BS_RWMF, BS_RED, BS_SNAP,
       BS_COUNT };

struct BatchArgs {
  const EngineArgs *args;  // [n] (device)
  const int *off;          // [BS_COUNT][n + 1] block offsets of each stage (device)
  int n;
};

// The engine of block blk of a stage and the block's index within it: the
// last engine whose first block is at or before blk (engines with no block
// in the stage have equal offsets and are passed over).
__device__ __forceinline__ int batch_find(const BatchArgs &ba, int stage, int blk, int &local) {
  const int *o = ba.off + stage * (ba.n + 1);
  int lo = 0, hi = ba.n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (o[mid] <= blk) lo = mid;
    else hi = mid;
  }
  local = blk - o[lo];
  return lo;
}

template <bool REC, int SHAPE, int DS>
__global__ void __launch_bounds__(PHX_BLOCK) k_batch_pass(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, BS_PASS, blockIdx.x, loc);
  const EngineArgs &a = ba.args[b];
  const phx_engine &e = a.e;
  double *const *c = e.coords;
  if (e.kick_shape == PHX_KICKS_DXDY)
    pass_body<0, REC, true, SHAPE, PHX_KICKS_DXDY, DS>(a, c[0], c[1], c[2], c[3], c[4], c[5],
                                                       e.programs, e.kicks, loc, a.grid);
  else if (e.kick_shape == PHX_KICKS_PLATES)
    pass_body<0, REC, true, SHAPE, PHX_KICKS_PLATES, DS>(a, c[0], c[1], c[2], c[3], c[4], c[5],
                                                         e.programs, e.kicks, loc, a.grid);
  else
    pass_body<0, REC, true, SHAPE, PHX_KICKS_GENERIC, DS>(a, c[0], c[1], c[2], c[3], c[4], c[5],
                                                          e.programs, e.kicks, loc, a.grid);
}

template <bool REC>
__global__ void __launch_bounds__(EN_HIST_THREADS) k_batch_cuts(BatchArgs ba) {
  __shared__ double res[PassLayout<REC, true>::NV];
  __shared__ phx_slice_state st;
  cuts_body<REC, true>(ba.args[blockIdx.x], res, &st);
}

template <bool REC>
__global__ void __launch_bounds__(EN_HIST_THREADS) k_batch_hist(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, BS_HIST, blockIdx.x, loc);
  const EngineArgs &a = ba.args[b];
  hist_body<REC, false>(a, loc, hist_blocks(a.e.n));
}

__global__ void __launch_bounds__(PHX_BLOCK) k_batch_merge(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, BS_MERGE, blockIdx.x, loc);
  const EngineArgs &a = ba.args[b];
  if (merge_sl(a.e.n_slices) == 4) merge_body<4>(a, loc);
  else merge_body<8>(a, loc);
}

__global__ void __launch_bounds__(WK_THREADS) k_batch_conv(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, BS_CONV, blockIdx.x, loc);
  const EngineArgs &a = ba.args[b];
  const int tiles = (a.e.n_slices + WK_TILE - 1) / WK_TILE;
  const int tile = loc % tiles, rest = loc / tiles;
  conv_body(a, tile, rest % a.n_kicks, rest / a.n_kicks);
}

template <int KM>
__global__ void __launch_bounds__(rw_threads<KM>()) k_batch_conv_rw(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, KM == 0 ? BS_RW : BS_RWMF, blockIdx.x, loc);
  conv_rw_body<KM>(ba.args[b], loc);
}

__global__ void k_batch_conv_reduce(BatchArgs ba) {
  int loc;
  const int b = batch_find(ba, BS_RED, blockIdx.x, loc);
  conv_reduce_body(ba.args[b], loc);
}

// Particle 0 of every engine: the shifts of the next pass.
__global__ void k_batch_snapshot(BatchArgs ba) {
  const EngineArgs &a = ba.args[blockIdx.x];
  EngineWs w = engine_ws(a.e.ws, a.e.n, a.e.n_slices);
  if (threadIdx.x < 6) w.p0[threadIdx.x] = a.e.coords[threadIdx.x][0];
}

struct Batch {
  int n;
  int rec, shape, dets;
  int total[BS_COUNT];
  size_t lds_pass, lds_hist, lds_conv, lds_rw;
  void *desc;
  std::vector<char> host;  // the descriptor image, kept for the copy
  BatchArgs args() const {
    BatchArgs b;
    b.args = reinterpret_cast<const EngineArgs *>(desc);
    b.off = reinterpret_cast<const int *>(reinterpret_cast<const char *>(desc) +
                                          round256((size_t)n * sizeof(EngineArgs)));
    b.n = n;
    return b;
  }
};

size_t batch_desc_bytes(int n) {
  return round256((size_t)n * sizeof(EngineArgs)) + (size_t)BS_COUNT * (n + 1) * sizeof(int);
}

// Blocks of each stage for one engine.
void stage_blocks(const EngineArgs &a, int (&nb)[BS_COUNT]) {
  const phx_engine &e = a.e;
  const int S = e.n_slices, ck = e.conv_kinds;
  const int rw_tiles = (S + RW_TILE - 1) / RW_TILE;
  nb[BS_PASS] = a.grid;
  nb[BS_CUTS] = 1;
  nb[BS_HIST] = hist_blocks(e.n);
  const bool has_wake = S > 0 && a.n_kicks > 0;
  nb[BS_MERGE] = (S + merge_sl(S) - 1) / merge_sl(S) + (has_wake && engine_has_rw(e) ? 1 : 0);
  nb[BS_CONV] = (ck == 0 || (ck & PHX_CONV_GENERAL)) ? (S + WK_TILE - 1) / WK_TILE * a.n_kicks * a.T : 0;
  nb[BS_RW] = (ck == 0 || (ck & PHX_CONV_RW_DOT)) ? rw_tiles * conv_rows(S, a.T) : 0;
  nb[BS_RWMF] = (ck == 0 || (ck & PHX_CONV_RW_MFMA)) ? rw_tiles * conv_rows(S, a.T) : 0;
  nb[BS_RED] = conv_reduces(S, a.T, engine_has_rw(e)) ? (a.n_kicks * S * RED_LANES + 255) / 256 : 0;
  nb[BS_SNAP] = 1;
}

// The compiled batch passes: first-order chromaticity (with or without
// amplitude detuning; general or smooth segments) after an RF system or a
// linear map, and the generic interpreter for every other map (and for
// statistics records).
template <bool REC, int SH, int DS>
void launch_batch_pass(const Batch &B, hipStream_t s) {
  k_batch_pass<REC, SH, DS><<<dim3((unsigned)B.total[BS_PASS]), PHX_BLOCK, B.lds_pass, s>>>(
      B.args());
}

template <int SH>
bool batch_pass_dets(const Batch &B, hipStream_t s) {
  switch (B.dets) {
    case PHX_DETS_C1: launch_batch_pass<false, SH, PHX_DETS_C1>(B, s); return true;
    case PHX_DETS_C1_AMP: launch_batch_pass<false, SH, PHX_DETS_C1_AMP>(B, s); return true;
    case PHX_DETS_C1 | PHX_DETS_SMOOTH:
      launch_batch_pass<false, SH, PHX_DETS_C1 | PHX_DETS_SMOOTH>(B, s);
      return true;
    case PHX_DETS_C1_AMP | PHX_DETS_SMOOTH:
      launch_batch_pass<false, SH, PHX_DETS_C1_AMP | PHX_DETS_SMOOTH>(B, s);
      return true;
    default: return false;
  }
}

int batch_pass(const Batch &B, hipStream_t s) {
  bool done = false;
  if (B.rec) {
    launch_batch_pass<true, PHX_SHAPE_GENERIC, PHX_DETS_GENERIC>(B, s);
    done = true;
  } else if (B.shape == PHX_SHAPE_RF1_T) {
    done = batch_pass_dets<PHX_SHAPE_RF1_T>(B, s);
  } else if (B.shape == PHX_SHAPE_LIN_T) {
    done = batch_pass_dets<PHX_SHAPE_LIN_T>(B, s);
  }
  if (!done) launch_batch_pass<false, PHX_SHAPE_GENERIC, PHX_DETS_GENERIC>(B, s);
  PHX_CHECK_LAUNCH("phx_batch: pass");
  return PHX_OK;
}

int batch_turn(const Batch &B, hipStream_t s) {
  const BatchArgs ba = B.args();
  int rc = batch_pass(B, s);
  if (rc) return rc;
  if (B.rec) k_batch_cuts<true><<<dim3(B.n), EN_HIST_THREADS, 0, s>>>(ba);
  else k_batch_cuts<false><<<dim3(B.n), EN_HIST_THREADS, 0, s>>>(ba);
  PHX_CHECK_LAUNCH("phx_batch: cuts");
  if (B.rec) k_batch_hist<true><<<dim3(B.total[BS_HIST]), EN_HIST_THREADS, B.lds_hist, s>>>(ba);
  else k_batch_hist<false><<<dim3(B.total[BS_HIST]), EN_HIST_THREADS, B.lds_hist, s>>>(ba);
  PHX_CHECK_LAUNCH("phx_batch: histogram");
  k_batch_merge<<<dim3(B.total[BS_MERGE]), PHX_BLOCK, 0, s>>>(ba);
  PHX_CHECK_LAUNCH("phx_batch: merge");
  if (B.total[BS_CONV]) {
    k_batch_conv<<<dim3(B.total[BS_CONV]), WK_THREADS, B.lds_conv, s>>>(ba);
    PHX_CHECK_LAUNCH("phx_batch: convolution");
  }
  if (B.total[BS_RW]) {
    k_batch_conv_rw<0><<<dim3(B.total[BS_RW]), rw_threads<0>(), B.lds_rw, s>>>(ba);
    PHX_CHECK_LAUNCH("phx_batch: resistive wall");
  }
  if (B.total[BS_RWMF]) {
    k_batch_conv_rw<1><<<dim3(B.total[BS_RWMF]), rw_threads<1>(), B.lds_rw, s>>>(ba);
    PHX_CHECK_LAUNCH("phx_batch: resistive wall (matrix cores)");
  }
  if (B.total[BS_RED]) {
    k_batch_conv_reduce<<<dim3(B.total[BS_RED]), 256, 0, s>>>(ba);
    PHX_CHECK_LAUNCH("phx_batch: sum over turns");
  }
  return PHX_OK;
}

int batch_turns(const Batch &B, int n_turns, hipStream_t s, bool snapshot) {
  if (snapshot && n_turns > 0) {
    k_batch_snapshot<<<dim3(B.n), 64, 0, s>>>(B.args());
    PHX_CHECK_LAUNCH("phx_batch: snapshot");
  }
  for (int t = 0; t < n_turns; ++t) {
    const int rc = batch_turn(B, s);
    if (rc) return rc;
  }
  return PHX_OK;
}

}  // namespace

extern "C" int64_t phx_batch_desc_bytes(int32_t n) {
  if (n < 1) return -1;
  return (int64_t)batch_desc_bytes(n);
}

extern "C" int phx_batch_create(const phx_engine *engines, int32_t n, void *desc,
                                int64_t desc_bytes, phx_stream_t stream, void **batch) {
  if (!batch) return phx::fail(PHX_E_INVALID, "phx_batch_create: null output");
  *batch = nullptr;
  if (!engines || n < 1) return phx::fail(PHX_E_INVALID, "phx_batch_create: no engines");
  if (!desc || desc_bytes < (int64_t)batch_desc_bytes(n))
    return phx::fail(PHX_E_INVALID, "phx_batch_create: descriptor buffer of %lld bytes, %lld "
                     "needed", (long long)desc_bytes, (long long)batch_desc_bytes(n));
  Batch *B = new Batch();
  B->n = n;
  B->desc = desc;
  B->host.assign(batch_desc_bytes(n), 0);
  EngineArgs *args = reinterpret_cast<EngineArgs *>(B->host.data());
  int *off = reinterpret_cast<int *>(B->host.data() + round256((size_t)n * sizeof(EngineArgs)));
  for (int k = 0; k < BS_COUNT; ++k) B->total[k] = 0;
  B->lds_hist = B->lds_conv = B->lds_rw = 0;
  for (int i = 0; i < n; ++i) {
    EngineArgs a;
    int rc = engine_args_for(&engines[i], &a);
    if (rc) {
      delete B;
      return rc;
    }
    const phx_engine &e = a.e;
    if (e.n_slices < 1 || a.n_kicks < 1) {
      delete B;
      return phx::fail(PHX_E_INVALID, "phx_batch_create: engine %d has no wake field", i);
    }
    if (e.sc_stages != 0) {
      delete B;
      return phx::fail(PHX_E_INVALID,
                       "phx_batch_create: engine %d has space-charge stages (not batched)", i);
    }
    // the batch kernels run the general RF code for stationary sequences
    const int shape = e.program_shape & ~PHX_SHAPE_STATIONARY;
    if (i == 0) {
      B->rec = e.record_stats;
      B->shape = shape;
      B->dets = e.det_shape;
    } else if (e.record_stats != B->rec || shape != B->shape || e.det_shape != B->dets) {
      delete B;
      return phx::fail(PHX_E_INVALID, "phx_batch_create: engine %d differs from engine 0 in "
                       "record_stats, program_shape or det_shape", i);
    }
    args[i] = a;
    int nb[BS_COUNT];
    stage_blocks(a, nb);
    for (int k = 0; k < BS_COUNT; ++k) {
      off[k * (n + 1) + i] = B->total[k];
      if ((int64_t)B->total[k] + nb[k] > 0x7fffffff) {
        delete B;
        return phx::fail(PHX_E_LIMIT, "phx_batch_create: too many blocks in stage %d", k);
      }
      B->total[k] += nb[k];
    }
    const size_t lh = acc_lds_bytes(e.n_slices, 2);
    B->lds_hist = lh > B->lds_hist ? lh : B->lds_hist;
    if (nb[BS_CONV]) {
      const size_t lc = wake_conv_lds_bytes(e.n_slices, a.gmax);
      B->lds_conv = lc > B->lds_conv ? lc : B->lds_conv;
    }
    if (nb[BS_RW] || nb[BS_RWMF]) {
      const size_t lr = wake_rw_lds_bytes(e.n_slices);
      B->lds_rw = lr > B->lds_rw ? lr : B->lds_rw;
    }
  }
  for (int k = 0; k < BS_COUNT; ++k) off[k * (n + 1) + n] = B->total[k];
  constexpr int nv_rec = PassLayout<true, true>::NV, nv = PassLayout<false, true>::NV;
  B->lds_pass = (B->rec ? BR_LDS(nv_rec) : BR_LDS(nv)) * sizeof(double);
  const hipError_t err = hipMemcpyAsync(desc, B->host.data(), B->host.size(),
                                        hipMemcpyHostToDevice, phx::as_stream(stream));
  if (err != hipSuccess) {
    delete B;
    return phx::fail((int)err, "phx_batch_create: descriptor copy: %s", hipGetErrorString(err));
  }
  *batch = B;
  return PHX_OK;
}

extern "C" int phx_batch_turns(void *batch, int32_t n_turns, phx_stream_t stream) {
  if (!batch) return phx::fail(PHX_E_INVALID, "phx_batch_turns: null batch");
  if (n_turns < 0) return phx::fail(PHX_E_INVALID, "phx_batch_turns: n_turns < 0");
  return batch_turns(*reinterpret_cast<Batch *>(batch), n_turns, phx::as_stream(stream), true);
}

extern "C" int phx_batch_snapshot(void *batch, phx_stream_t stream) {
  if (!batch) return phx::fail(PHX_E_INVALID, "phx_batch_snapshot: null batch");
  const Batch &B = *reinterpret_cast<Batch *>(batch);
  k_batch_snapshot<<<dim3(B.n), 64, 0, phx::as_stream(stream)>>>(B.args());
  PHX_CHECK_LAUNCH("phx_batch_snapshot");
  return PHX_OK;
}

extern "C" int phx_batch_capture(void *batch, int32_t n_turns, void **graph) {
  if (!graph) return phx::fail(PHX_E_INVALID, "phx_batch_capture: null output");
  *graph = nullptr;
  if (!batch) return phx::fail(PHX_E_INVALID, "phx_batch_capture: null batch");
  if (n_turns < 1) return phx::fail(PHX_E_INVALID, "phx_batch_capture: n_turns < 1");
  const Batch &B = *reinterpret_cast<Batch *>(batch);
  hipStream_t cs;
  PHX_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "phx_batch_capture");
  hipGraph_t g = nullptr;
  hipError_t err = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    (void)hipStreamDestroy(cs);
    return phx::fail((int)err, "phx_batch_capture: begin: %s", hipGetErrorString(err));
  }
  const int rc = batch_turns(B, n_turns, cs, false);
  err = hipStreamEndCapture(cs, &g);
  (void)hipStreamDestroy(cs);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (err != hipSuccess) return phx::fail((int)err, "phx_batch_capture: end: %s", hipGetErrorString(err));
  hipGraphExec_t ex = nullptr;
  err = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (err != hipSuccess)
    return phx::fail((int)err, "phx_batch_capture: instantiate: %s", hipGetErrorString(err));
  *graph = new phx::GraphHandle{ex};
  return PHX_OK;
}

extern "C" int phx_batch_destroy(void *batch) {
  delete reinterpret_cast<Batch *>(batch);
  return PHX_OK;
}
