// Device side of the fused one-turn map (the engine): workspace layout,
// launch arguments and the bodies of the per-turn stages, shared by the
// single-engine kernels (phx_engine.hip) and the batch kernels, which run
// one stage of many engines in one launch (phx_batch.hip).
#pragma once

#include "phx_device.h"
#include "phx_sc.h"
#include "phx_slicing.h"
#include "phx_wake.h"

#define EN_STAT_NV 33
#define EN_NV (ZB_NV + EN_STAT_NV)
// the slice histogram: at most EN_HIST_BLOCKS blocks (partial histograms)
// of EN_HIST_THREADS threads, EN_HIST_PPT particles per thread and chunk
#ifndef EN_HIST_BLOCKS
#define EN_HIST_BLOCKS 256
#endif
#ifndef EN_HIST_THREADS
#define EN_HIST_THREADS 1024
#endif
#ifndef EN_HIST_PPT
#define EN_HIST_PPT PHX_PPT
#endif

namespace phx_en {


inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

struct EngineWs {
  double *partials;   // [EN_NV][pass grid] (slot-major)
  double *p0;         // particle 0 before the pass: the shifts of its sums
  AccParts acc;       // H partial histograms, T = 2 (x, y)
  double *conv_part;  // [n_kicks][conv_rows(S, T)][S]
};

__host__ __device__ inline int hist_blocks(int64_t n) {
  const int64_t nc = (n + PHX_CHUNK - 1) / PHX_CHUNK;
  return (int)(nc < EN_HIST_BLOCKS ? nc : EN_HIST_BLOCKS);
}

// The partials region holds the largest pass grid, so the layout does not
// depend on the grid chosen.
__host__ __device__ inline EngineWs engine_ws(void *ws, int64_t n, int S) {
  EngineWs w;
  char *p = reinterpret_cast<char *>(ws);
  const int64_t nc = phx_grid_for(n, PHX_MAX_GRID);
  const int Se = S > 0 ? S : 1;
  w.partials = reinterpret_cast<double *>(p);
  p += ((size_t)nc * EN_NV * 8 + 255) / 256 * 256;
  w.p0 = reinterpret_cast<double *>(p);
  p += 256;
  const int H = hist_blocks(n);
  w.acc = acc_parts_layout(p, H, Se, 2);
  p += acc_parts_bytes(H, Se, 2);
  w.conv_part = reinterpret_cast<double *>(p);
  return w;
}

// Whether the engine's convolution may have resistive-wall kicks in the
// dot-product form (conv_kinds 0: every kind).
__host__ __device__ inline bool engine_has_rw(const phx_engine &e) {
  return e.conv_kinds == 0 || (e.conv_kinds & (PHX_CONV_RW_DOT | PHX_CONV_RW_MFMA));
}

// The resistive-wall kick description (rw_desc_build), after the partial
// sums of the convolution.
__host__ __device__ inline RwKick *engine_rw_desc(const EngineWs &w, int n_kicks, int S, int T) {
  const int Se = S > 0 ? S : 1;
  return reinterpret_cast<RwKick *>(w.conv_part +
                                    (size_t)(n_kicks > 0 ? n_kicks : 1) * conv_rows(Se, T) * Se);
}

inline size_t engine_ws_bytes(int64_t n, int S, int n_kicks, int T) {
  const int64_t nc = phx_grid_for(n, PHX_MAX_GRID);
  const int Se = S > 0 ? S : 1;
  size_t b = round256((size_t)nc * EN_NV * 8) + 256 + acc_parts_bytes(hist_blocks(n), Se, 2);
  b += (size_t)(n_kicks > 0 ? n_kicks : 1) * conv_rows(Se, T) * Se * 8;
  b += rw_desc_bytes(T);
  return round256(b);
}

struct EngineArgs {
  phx_engine e;
  int n_kicks;
  int T;
  int grid;  // blocks of the particle pass (its partials: [slot][grid])
  int gmax;  // kicks one convolution block serves (LDS permitting)
};

// ---------------------------------------------------------------------------
// Space-charge stages.  Their workspace (sc_ws): the Faddeeva coefficients,
// the cuts of the last space-charge slicing, the field parameters of every
// slice (one 64-byte row each) and the partial histograms of counts and
// sums of x, y, x^2, y^2 (laid out for the slicing's own slice count).
// ---------------------------------------------------------------------------
#define SC_TERMS 4

struct ScSlice {
  double mx, my, nl, pad;
  ScField f;
};

struct ScWs {
  WofzCoeffs *wc;
  phx_slice_state *state;
  ScSlice *slices;
  char *acc;  // acc_parts_layout(acc, hist_blocks(n), S, SC_TERMS)
};

__host__ __device__ inline size_t sc_r256(size_t b) { return (b + 255) / 256 * 256; }

__host__ __device__ inline ScWs sc_ws(void *base, int S) {
  ScWs w;
  char *p = reinterpret_cast<char *>(base);
  w.wc = reinterpret_cast<WofzCoeffs *>(p);
  p += sc_r256(sizeof(WofzCoeffs));
  w.state = reinterpret_cast<phx_slice_state *>(p);
  p += sc_r256(sizeof(phx_slice_state));
  w.slices = reinterpret_cast<ScSlice *>(p);
  p += sc_r256((size_t)(S > 0 ? S : 1) * sizeof(ScSlice));
  w.acc = p;
  return w;
}

inline size_t sc_ws_bytes(int64_t n, int S) {
  const int Se = S > 0 ? S : 1;
  return sc_r256(sizeof(WofzCoeffs)) + sc_r256(sizeof(phx_slice_state)) +
         sc_r256((size_t)Se * sizeof(ScSlice)) + sc_r256(acc_parts_bytes(hist_blocks(n), Se, SC_TERMS));
}

// What a space-charge kick of a pass reads: the cuts of the slicing after
// the previous stage, its slices' field parameters, the coefficients.
struct ScCtx {
  const ScSlice *sl;
  const WofzCoeffs *wc;
  double tail, width, inv_width;
  int S;
};

template <int N>
__device__ __forceinline__ void sc_kick_n(P6 (&p)[N], const ScCtx &c, const phx_spacecharge &sp) {
  double dx[N], dy[N], nl[N], ex[N], ey[N];
  ScField f[N];
  bool in[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double fl = slice_floor_r(p[k].z, c.tail, c.width, c.inv_width);
    in[k] = fl >= 0.0 && fl < (double)c.S;
    const ScSlice &q = c.sl[in[k] ? (int)fl : 0];
    dx[k] = p[k].x - q.mx;
    dy[k] = p[k].y - q.my;
    nl[k] = q.nl;
    f[k] = q.f;
  }
  sc_field_n<N>(*c.wc, sp, dx, dy, f, ex, ey);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (in[k]) {
      p[k].xp = p[k].xp + (ex[k] * nl[k]) * sp.prefactor;
      p[k].yp = p[k].yp + (ey[k] * nl[k]) * sp.prefactor;
    }
  }
}

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

// nk = min(program kicks, engine kicks): the kick table holds the latter.
template <int KS>
__device__ __forceinline__ void kick_pending(P6 &p, const phx_turn_program &g, int nk,
                                             double tail, double width, double inv_width,
                                             int S, const double *kicks) {
  const double f = slice_floor_r(p.z, tail, width, inv_width);
  if (f >= 0.0 && f < (double)S) apply_kick_set<KS>(p, (int32_t)f, S, g.kick, nk, kicks);
}

// The pending kick's table values for the PHX_PPT particles of a group,
// read one group ahead of their use (the loads are otherwise a dependent
// gather in the middle of each group's arithmetic).  For the kick sets
// with straight-line code: DXDY (dipole x, y: 2 values per particle) and
// PLATES (dipole and quadrupole x, y: 4).  The kicks do not move z, so the
// slice index from the group's loaded z is the one kick_pending takes, and
// apply() adds in apply_kick_set's order: the same arithmetic.
template <int KS>
struct KickAhead {
  static constexpr int NV = KS == PHX_KICKS_DXDY ? 2 : (KS == PHX_KICKS_PLATES ? 4 : 0);
  static constexpr bool ON = NV > 0;
  double k[PHX_PPT][NV > 0 ? NV : 1];
  bool in[PHX_PPT];
  __device__ __forceinline__ void fetch(const P6 (&q)[PHX_PPT], double tail, double width,
                                        double inv_width, int S, const double *kicks) {
#pragma unroll
    for (int j = 0; j < PHX_PPT; ++j) {
      const double f = slice_floor_r(q[j].z, tail, width, inv_width);
      in[j] = f >= 0.0 && f < (double)S;
      const int s = in[j] ? (int)f : 0;  // slot 0 is read and not used outside
#pragma unroll
      for (int i = 0; i < NV; ++i) k[j][i] = kicks[(size_t)i * S + s];
    }
  }
  __device__ __forceinline__ void apply(P6 (&q)[PHX_PPT]) const {
#pragma unroll
    for (int j = 0; j < PHX_PPT; ++j) {
      if (!in[j]) continue;
      if constexpr (KS == PHX_KICKS_DXDY) {
        q[j].xp = q[j].xp + k[j][0];
        q[j].yp = q[j].yp + k[j][1];
      } else if constexpr (KS == PHX_KICKS_PLATES) {
        q[j].xp = q[j].xp + k[j][0];
        q[j].xp = q[j].xp + k[j][1] * q[j].x;
        q[j].yp = q[j].yp + k[j][2];
        q[j].yp = q[j].yp + k[j][3] * q[j].y;
      }
    }
  }
};

// The six coordinate arrays of a pass.
struct GroupPtr {
  double *c[6];
};

// PHX_PASS_NT: the momenta xp, yp, dp (which the slice histogram after the
// pass does not read) are loaded and stored non-temporally, so that they do
// not push x, y and z out of the last-level cache before the histogram;
// with PHX_PASS_NT_ALL every coordinate's (old) value is loaded so.
#ifndef PHX_PASS_NT
#define PHX_PASS_NT 1
#endif
#ifndef PHX_PASS_NT_ALL
#define PHX_PASS_NT_ALL 0
#endif
__device__ __forceinline__ double ld_mom(const double *p) {
#if PHX_PASS_NT
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ double ld_pos(const double *p) {
#if PHX_PASS_NT_ALL
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}
__device__ __forceinline__ void st_mom(double *p, double v) {
#if PHX_PASS_NT
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// Coordinates addressed as the array's base (the kernel arguments, kept in
// scalar registers) plus one 32-bit byte offset per particle, shared by the
// six arrays: per-group pointers to six arrays, current and next group,
// would take 24 scalar registers across the particle loop (bunches up to
// PHX_ENGINE_MAX_N keep every offset below 2^32).
__device__ __forceinline__ const double *at_off(const double *b, uint32_t off) {
  return reinterpret_cast<const double *>(reinterpret_cast<const char *>(b) + off);
}
__device__ __forceinline__ double *at_off(double *b, uint32_t off) {
  return reinterpret_cast<double *>(reinterpret_cast<char *>(b) + off);
}

// The PHX_PPT particles of this thread in the group starting at byte offset
// goff; lim: the group's particle count if it is the partial last group,
// else PHX_CHUNK (no guard needed).
__device__ __forceinline__ void group_load(const GroupPtr &gb, uint32_t goff, int lim,
                                           P6 (&p)[PHX_PPT]) {
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int o = (int)threadIdx.x + k * PHX_BLOCK;
    const uint32_t off = goff + (uint32_t)o * 8u;
    if (lim == PHX_CHUNK || o < lim)
      p[k] = P6{ld_pos(at_off(gb.c[0], off)), ld_mom(at_off(gb.c[1], off)),
                ld_pos(at_off(gb.c[2], off)), ld_mom(at_off(gb.c[3], off)),
                ld_pos(at_off(gb.c[4], off)), ld_mom(at_off(gb.c[5], off))};
    else
      p[k] = P6{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  }
}

__device__ __forceinline__ void group_store(const GroupPtr &gb, uint32_t goff, int lim,
                                            const P6 (&p)[PHX_PPT]) {
#pragma unroll
  for (int k = 0; k < PHX_PPT; ++k) {
    const int o = (int)threadIdx.x + k * PHX_BLOCK;
    const uint32_t off = goff + (uint32_t)o * 8u;
    if (lim == PHX_CHUNK || o < lim) {
      *at_off(gb.c[0], off) = p[k].x; st_mom(at_off(gb.c[1], off), p[k].xp);
      *at_off(gb.c[2], off) = p[k].y; st_mom(at_off(gb.c[3], off), p[k].yp);
      *at_off(gb.c[4], off) = p[k].z; st_mom(at_off(gb.c[5], off), p[k].dp);
    }
  }
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

template <bool REC, bool WAKE>
struct PassLayout {
  static constexpr int NV = (WAKE ? ZB_NV : 0) + (REC ? EN_STAT_NV : 0) + ((!WAKE && !REC) ? 1 : 0);
  static constexpr unsigned long long MINM = WAKE ? ZB_MINM : 0ULL;
  static constexpr unsigned long long MAXM = WAKE ? ZB_MAXM : 0ULL;
  static constexpr int SOFF = WAKE ? ZB_NV : 0;
};

// MODE 0: a turn; MODE 1: flush (pending kick + final statistics only).
// Persistent grid (a.grid blocks): each block walks its particle
// groups with the next group's coordinates loaded while the current one is
// computed, so the particle traffic overlaps the FP64 work.
// No workgroup barrier or fence precedes the last read of the turn program
// and the coordinate arrays are declared non-aliasing (__restrict__
// arguments): otherwise the compiler must assume the program may have been
// written and reloads it with vector loads that wait on the particle loads.
// So particle 0 is computed by every lane, and the kick table is read from
// global memory (it stays in the caches).  Dynamic LDS: the block reduction.
#ifndef PHX_PASS_MIN_BLOCKS
#define PHX_PASS_MIN_BLOCKS 1
#endif
#ifndef PHX_PASS_LAUNDER
#define PHX_PASS_LAUNDER 0
#endif
// Space-charge engines (sc_stages > 0) run this pass for stage 0 of a
// turn (stage: the offset of the stage's program from the turn's first;
// the later stages run sc_pass_body).  The pending wake kick and the
// statistics belong to stage 0.
template <int MODE, bool REC, bool WAKE, int SHAPE, int KS, int DS>
__device__ __forceinline__ void pass_body(
    const EngineArgs &a, double *__restrict__ x, double *__restrict__ xp,
    double *__restrict__ y, double *__restrict__ yp, double *__restrict__ z,
    double *__restrict__ dp, const phx_turn_program *__restrict__ programs,
    const double *__restrict__ kicks, int bid, int nblk, int stage = 0) {
  using L = PassLayout<REC, WAKE>;
  const phx_engine &e = a.e;
  extern __shared__ double pass_lds[];
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int64_t n = e.n;
  const int turn = *e.turn_counter;
  const int prog_turn = MODE == 0 ? turn : (turn > 0 ? turn - 1 : 0);
  const phx_turn_program &g = programs[e.turn_program[prog_turn] + stage];
  const int pend = stage == 0 ? *e.pending : 0;
  const double tail = e.state->tail, width = e.state->width;
  const double inv_width = 1.0 / width;
  const int S = e.n_slices;
  const int nk = g.n_kicks < a.n_kicks ? g.n_kicks : a.n_kicks;
  const int64_t ng = phx_n_groups(n);
  // the group's base pointers and particle count (uniform)
  const GroupPtr gb{{x, xp, y, yp, z, dp}};
  // the group's byte offset and particle count (uniform)
  auto group = [&](int64_t gg, uint32_t &off) {
    const int64_t b = gg * PHX_CHUNK;
    off = (uint32_t)b * 8u;
    const int64_t rest = n - b;
    return rest >= PHX_CHUNK ? PHX_CHUNK : (int)rest;
  };
  int64_t gi = bid;
  uint32_t gc, gnx = 0;
  int lc = group(gi, gc);
  P6 cur[PHX_PPT], nxt[PHX_PPT];
  group_load(gb, gc, lc, cur);
  // Particle 0 sets the shifts of the shifted sums (stats.pyx:cov): its
  // state after the pending kick (statistics), and its z after this turn's
  // elements (cuts), which needs the kick only where a kick changes dp and
  // only the longitudinal elements.  Every lane computes it (uniform values;
  // once per block) from the snapshot taken before this pass: block 0
  // rewrites particle 0 while later blocks may still be starting.
  double src[7];
  {
    const double *p0 = w.p0;
    P6 q = P6{p0[0], p0[1], p0[2], p0[3], p0[4], p0[5]};
    bool kick_dp = false;
    if constexpr (KS == PHX_KICKS_GENERIC) {
      for (int k = 0; k < nk; ++k) kick_dp = kick_dp || g.kick[k].plane == PHX_PLANE_DP;
    }
    if (pend && (REC || kick_dp)) kick_pending<KS>(q, g, nk, tail, width, inv_width, S, kicks);
    src[0] = q.x; src[1] = q.xp; src[2] = q.y; src[3] = q.yp; src[4] = q.z; src[5] = q.dp;
    if (MODE == 0) apply_shape_longitudinal<SHAPE>(q, g);
    src[6] = q.z;
  }
  double v[L::NV];
#pragma unroll
  for (int i = 0; i < L::NV; ++i) v[i] = red_ident<L::MINM, L::MAXM>(i);
  double shs[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) shs[i] = src[i];
  const double zshift = src[6];
  KickAhead<KS> ka;
  if (KickAhead<KS>::ON && pend) ka.fetch(cur, tail, width, inv_width, S, kicks);
  for (; gi < ng; gi += nblk) {
    const int64_t gn = gi + nblk;
    int ln = 0;
    if (gn < ng) {
      ln = group(gn, gnx);
      group_load(gb, gnx, ln, nxt);
    }
    if (pend) {
      if constexpr (KickAhead<KS>::ON) {
        ka.apply(cur);
      } else {
#pragma unroll
        for (int k = 0; k < PHX_PPT; ++k)
          kick_pending<KS>(cur[k], g, nk, tail, width, inv_width, S, kicks);
      }
    }
    const bool full = lc == PHX_CHUNK;
    if (REC) {
#pragma unroll
      for (int k = 0; k < PHX_PPT; ++k)
        if (full || (int)threadIdx.x + k * PHX_BLOCK < lc)
          stats_accumulate(&v[L::SOFF], cur[k], shs);
    }
    if (MODE == 0) {
      // PHX_PASS_LAUNDER: read the element parameters from the program in
      // every group (a pointer the compiler cannot follow across
      // iterations) instead of holding them in scalar registers over the
      // loop; measured slower (more vector registers, 2 waves per SIMD)
      const phx_turn_program *gl = &g;
#if PHX_PASS_LAUNDER
      asm volatile("" : "+s"(gl));
#endif
      apply_shape_n<SHAPE, PHX_PPT, DS>(cur, *gl);
    }
    if (WAKE) {
#pragma unroll
      for (int k = 0; k < PHX_PPT; ++k) {
        if (full || (int)threadIdx.x + k * PHX_BLOCK < lc) {
          double zb[ZB_NV];
#pragma unroll
          for (int i = 0; i < ZB_NV; ++i) zb[i] = v[i];
          zblock_accumulate<0x05u>(zb, cur[k], zshift);  // max |x|, |y|: the histogram's
#pragma unroll
          for (int i = 0; i < ZB_NV; ++i) v[i] = zb[i];
        }
      }
    }
    if (MODE == 0 || pend) group_store(gb, gc, lc, cur);
#pragma unroll
    for (int k = 0; k < PHX_PPT; ++k) cur[k] = nxt[k];
    gc = gnx;
    lc = ln;
    // the next group's kick values, while this block's other waves compute
    if (KickAhead<KS>::ON && pend && gn < ng) ka.fetch(cur, tail, width, inv_width, S, kicks);
  }
  const double r = block_reduce_slots<L::NV, L::MINM, L::MAXM>(v, pass_lds);
  if (threadIdx.x < L::NV) w.partials[(size_t)threadIdx.x * nblk + bid] = r;
}

// The pass of space-charge stage `stage` > 0: the space-charge kick that
// ended the previous stage (the slicing after it is in the space-charge
// workspace), then the stage's elements, the coordinates written back and
// the z block of the result (11 partials per block, as pass_body without
// statistics).  The field evaluation needs many registers, so a thread
// holds one particle at a time, taking the particles of pass_body's groups
// in pass_body's order: the z-block sums, and so the cuts, are those
// pass_body would give.
template <int SHAPE, int DS>
__device__ __forceinline__ void sc_pass_body(
    const EngineArgs &a, double *__restrict__ x, double *__restrict__ xp,
    double *__restrict__ y, double *__restrict__ yp, double *__restrict__ z,
    double *__restrict__ dp, const phx_turn_program *__restrict__ programs, int bid, int nblk,
    int stage) {
  const phx_engine &e = a.e;
  extern __shared__ double pass_lds[];
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int64_t n = e.n;
  const int turn = *e.turn_counter;
  const int first = e.turn_program[turn];
  const phx_turn_program &g = programs[first + stage];
  const phx_spacecharge &sp = programs[first + stage - 1].sc;
  const ScWs sw = sc_ws(e.sc_ws, e.sc_slices);
  ScCtx scc;
  scc.sl = sw.slices;
  scc.wc = sw.wc;
  scc.tail = sw.state->tail;
  scc.width = sw.state->width;
  scc.inv_width = 1.0 / scc.width;
  scc.S = sw.state->n_slices;
  // particle 0's z after this stage's elements: the shift of the z sums
  double zshift;
  {
    P6 q = P6{w.p0[0], w.p0[1], w.p0[2], w.p0[3], w.p0[4], w.p0[5]};
    apply_shape_longitudinal<SHAPE>(q, g);
    zshift = q.z;
  }
  // transverse-only stages leave z and dp as they are
  const bool store_long = SHAPE == PHX_SHAPE_T || SHAPE == PHX_SHAPE_NONE
                              ? false
                              : (SHAPE != PHX_SHAPE_GENERIC || g.n_long > 0 || g.n_lin > 0);
  double v[ZB_NV];
#pragma unroll
  for (int i = 0; i < ZB_NV; ++i) v[i] = red_ident<ZB_MINM, ZB_MAXM>(i);
  const int64_t ng = phx_n_groups(n);
  for (int64_t gi = bid; gi < ng; gi += nblk) {
    const int64_t b = gi * PHX_CHUNK;
    const int64_t rest = n - b;
    const int lim = rest >= PHX_CHUNK ? PHX_CHUNK : (int)rest;
#pragma unroll 1
    for (int k = 0; k < PHX_PPT; ++k) {
      const int o = (int)threadIdx.x + k * PHX_BLOCK;
      if (o < lim) {
        const int64_t j = b + o;
        P6 p[1];
        p[0] = P6{ld_pos(x + j), ld_mom(xp + j), ld_pos(y + j), ld_mom(yp + j), ld_pos(z + j),
                  ld_mom(dp + j)};
        sc_kick_n<1>(p, scc, sp);
        apply_shape_n<SHAPE, 1, DS>(p, g);
        x[j] = p[0].x;
        st_mom(xp + j, p[0].xp);
        y[j] = p[0].y;
        st_mom(yp + j, p[0].yp);
        if (store_long) {
          z[j] = p[0].z;
          st_mom(dp + j, p[0].dp);
        }
        zblock_accumulate<0x05u>(v, p[0], zshift);
      }
    }
  }
  const double r = block_reduce_slots<ZB_NV, ZB_MINM, ZB_MAXM>(v, pass_lds);
  if (threadIdx.x < ZB_NV) w.partials[(size_t)threadIdx.x * nblk + bid] = r;
}

// The cuts of this turn from the pass partials (one block): fold the z
// block (and the statistics slots) in a fixed order, the slicing state, the
// statistics row, particle 0 for the next pass, and the wake-history
// bookkeeping (WakeField.track: ages += C/(beta c), then the new entry at
// age 0, impedances/wakes.py:134-142).  FOLD_ALL: this block does every
// part (the batch's cuts stage, block 0 of the standalone histogram);
// otherwise only the fold into res and the shared cut values (the other
// histogram blocks).  Returns the state through *stp (thread 0).
template <bool REC, bool FOLD_ALL>
__device__ __forceinline__ void cuts_body(const EngineArgs &a, double *res,
                                          phx_slice_state *stp) {
  using L = PassLayout<REC, true>;
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int turn = *e.turn_counter;
  const phx_turn_program &g = e.programs[e.turn_program[turn]];
  if (FOLD_ALL && threadIdx.x < 6) w.p0[threadIdx.x] = e.coords[threadIdx.x][0];
  fold_partials<L::MINM, L::MAXM>(w.partials, a.grid, 0, ZB_NV, res);
  if constexpr (REC && FOLD_ALL) fold_partials<L::MINM, L::MAXM>(w.partials, a.grid, ZB_NV, L::NV, res);
  // the wake history: every stored entry ages by the turn's increment, one
  // thread per entry (a serial loop made the block wait on one round trip
  // per entry); the slot the new entry takes when the ring is full is
  // skipped, since thread 0 then sets it to age 0
  phx_wake_hist *h = e.hist;
  int cap = 1, nf = 0, nh = 0;
  if (FOLD_ALL) {
    const int head = h->head;
    cap = h->capacity;
    nf = h->n_filled;
    nh = (head + 1) % cap;
    for (int i = threadIdx.x; i < nf; i += blockDim.x) {
      const int sl = ring_slot(head, i, cap);
      if (sl != nh) h->age[sl] = h->age[sl] + g.age_increment;
    }
    lds_sync();  // every thread has read the ring's state before thread 0 moves it
  }
  if (threadIdx.x == 0) {
    phx_slice_state st;
    compute_cuts(g.slicer, e.n, res[0], res[1], res[2], res[3], res[4], st);
    for (int i = 0; i < 6; ++i) st.absmax[i] = res[5 + i];
    st.turn = turn;
    st.pad = 0;
    *stp = st;
    if (FOLD_ALL) {
      *e.state = st;
      h->head = nh;
      h->n_filled = nf + 1 < cap ? nf + 1 : cap;
      h->age[nh] = 0.0;
      h->beta[nh] = g.beta;
    }
  }
  if (REC && FOLD_ALL && threadIdx.x < EN_STAT_NV)
    e.stats[(size_t)turn * EN_STAT_NV + threadIdx.x] = res[L::SOFF + threadIdx.x];
}

// Slice histogram of the new state: counts, sum x and sum y per slice in an
// LDS histogram, stored as partial bid of nblk.  FOLD: the block first folds
// the pass partials into the cuts itself (every block, identical results;
// block 0 also publishes them and does the turn's bookkeeping); otherwise
// the cuts stage did that and the block reads the state.
template <bool REC, bool FOLD>
__device__ __forceinline__ void hist_body(const EngineArgs &a, int bid, int nblk) {
  using L = PassLayout<REC, true>;
  extern __shared__ unsigned long long acc_lds[];
  __shared__ double res[L::NV];
  __shared__ phx_slice_state sh_st;
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int S = e.n_slices;
  constexpr int CH = EN_HIST_THREADS * EN_HIST_PPT;
  const int64_t nc = (e.n + CH - 1) / CH;
  const int64_t n = e.n;
  // Each chunk loads the next one before it is binned.  No particle load is
  // in flight across a barrier: with the first chunk's loads issued ahead
  // of the fold of the cuts, some blocks' folds came out wrong on MI355X
  // (about 1 in 25 block folds on a new engine's first turns; none with
  // the loads issued after the last barrier).
  double zz[EN_HIST_PPT], xx[EN_HIST_PPT], yy[EN_HIST_PPT];
  double zn[EN_HIST_PPT], xn[EN_HIST_PPT], yn[EN_HIST_PPT];
  auto load_chunk = [&](int64_t c, double (&z)[EN_HIST_PPT], double (&x)[EN_HIST_PPT],
                        double (&y)[EN_HIST_PPT]) {
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      const int64_t j = c * CH + threadIdx.x + (int64_t)k * EN_HIST_THREADS;
      if (c < nc && j < n) {
        z[k] = e.coords[4][j];
        x[k] = e.coords[0][j];
        y[k] = e.coords[2][j];
      } else {
        z[k] = 0.0;
        x[k] = 0.0;
        y[k] = 0.0;
      }
    }
  };
  if constexpr (FOLD) {
    if (bid == 0) cuts_body<REC, true>(a, res, &sh_st);
    else cuts_body<REC, false>(a, res, &sh_st);
  } else {
    if (threadIdx.x == 0) sh_st = *e.state;
  }
  AccLds A = acc_lds_layout(acc_lds, S, 2);
  acc_lds_zero(A, S, 2);  // ends with a barrier: the shared cuts are visible
  const double tail = sh_st.tail, width = sh_st.width;
  const double inv_width = 1.0 / width;
  const int ex = fx_exponent(sh_st.absmax[0]), ey = fx_exponent(sh_st.absmax[2]);
  load_chunk(bid, zz, xx, yy);
  for (int64_t c = bid; c < nc; c += nblk) {
    load_chunk(c + nblk, zn, xn, yn);
    const int64_t base = c * CH + threadIdx.x;
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      const int64_t j = base + (int64_t)k * EN_HIST_THREADS;
      if (j < n) {
        const double f = slice_floor_r(zz[k], tail, width, inv_width);
        if (f >= 0.0 && f < (double)S) {
          const int s = (int)f;
          atomicAdd(&A.counts[s], 1u);
          acc_add_value(A, S, 0, s, xx[k], ex, w.acc.fallback);
          acc_add_value(A, S, 1, s, yy[k], ey, w.acc.fallback);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      zz[k] = zn[k];
      xx[k] = xn[k];
      yy[k] = yn[k];
    }
  }
  acc_lds_store(A, S, 2, w.acc, bid);
}

// Merge the partial histograms; means and the new history entry:
// times = z_centers / (beta c) (particles/slicing.py:124-126, 303-307),
// moments n, n*mean_x, n*mean_y (impedances/wake_kicks.py:217, 235).
// Block (S + SL - 1) / SL, launched when the engine has resistive-wall kicks
// in the dot-product form, describes the turn's kicks for the convolution
// (rw_desc_build) while the others merge.
template <int SL>
__device__ __forceinline__ void merge_body(const EngineArgs &a, int bid) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const int S = e.n_slices;
  SliceTotals tot;
  const int turn = e.state->turn;
  if (bid == (S + SL - 1) / SL) {
    if (a.n_kicks < 1 || !e.hist) return;
    const phx_turn_program &g = e.programs[e.turn_program[turn]];
    rw_desc_build(e.hist, *reinterpret_cast<const WakeKicks *>(g.kick), a.n_kicks, a.T,
                  engine_rw_desc(w, a.n_kicks, S, a.T));
    return;
  }
  if (bid == 0 && threadIdx.x == 0) *e.turn_counter = turn + 1;
  if (!acc_merge<2, SL>(w.acc, S, tot, bid)) return;
  const int s = bid * SL + threadIdx.x % SL;
  // bin edges as np.arange(tail, head + 1e-7*width, width)
  const double tail = e.state->tail, width = e.state->width;
  e.z_bins[s] = arange_value(tail, width, s);
  if (s == S - 1) e.z_bins[S] = arange_value(tail, width, S);
  const phx_turn_program &g = e.programs[e.turn_program[turn]];
  const int ex = fx_exponent(e.state->absmax[0]);
  const int ey = fx_exponent(e.state->absmax[2]);
  const double cnt = (double)tot.count;
  const double sx = fx_value(tot.hi[0], tot.lo[0], ex) + w.acc.fallback[s];
  const double sy = fx_value(tot.hi[1], tot.lo[1], ey) + w.acc.fallback[S + s];
  w.acc.fallback[s] = 0.0;
  w.acc.fallback[S + s] = 0.0;
  const double mx = tot.count ? sx / cnt : 0.0;
  const double my = tot.count ? sy / cnt : 0.0;
  e.counts[s] = (int32_t)tot.count;
  e.slice_means[s] = mx;
  e.slice_means[S + s] = my;
  phx_wake_hist *h = e.hist;
  const int nh = h->head;
  const double zl = arange_value(tail, width, s), zr = arange_value(tail, width, s + 1);
  const double zc = zl + 0.5 * (zr - zl);
  h->times[(size_t)nh * S + s] = zc / g.beta_c;
  double *mm = h->mom + (size_t)nh * 3 * S;
  mm[s] = cnt;
  mm[S + s] = cnt * mx;
  mm[2 * S + s] = cnt * my;
}

// Space-charge slicing of the state after stage `stage` (the slicing that
// TransverseGaussianSpaceCharge.track asks for with mean_x, mean_y, sigma_x,
// sigma_y, spacecharge/spacecharge.py:207-208): every block folds the
// stage pass's z block into the cuts of the stage program's sc_slicer
// (block 0 publishes them and records particle 0 for the next pass), then
// keeps an LDS histogram of counts and fixed-point sums of x, y, x^2 and
// y^2, stored as its partial.  The scales come from max |x| and max |y| of
// the same state, so every term fits.
__device__ __forceinline__ void sc_hist_body(const EngineArgs &a, int stage, int bid, int nblk) {
  extern __shared__ unsigned long long acc_lds[];
  __shared__ double res[ZB_NV];
  __shared__ phx_slice_state sh_st;
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const ScWs sw = sc_ws(e.sc_ws, e.sc_slices);
  const int turn = *e.turn_counter;
  const phx_turn_program &g = e.programs[e.turn_program[turn] + stage];
  const int S = g.sc_slicer.n_slices;
  const int64_t n = e.n;
  const AccParts P = acc_parts_layout(sw.acc, hist_blocks(n), S, SC_TERMS);
  fold_partials<ZB_MINM, ZB_MAXM>(w.partials, a.grid, 0, ZB_NV, res);
  if (threadIdx.x == 0) {
    phx_slice_state st;
    compute_cuts(g.sc_slicer, n, res[0], res[1], res[2], res[3], res[4], st);
    for (int i = 0; i < 6; ++i) st.absmax[i] = res[5 + i];
    st.turn = turn;
    st.pad = 0;
    sh_st = st;
    if (bid == 0) *sw.state = st;
  }
  if (bid == 0 && threadIdx.x < 6) w.p0[threadIdx.x] = e.coords[threadIdx.x][0];
  AccLds A = acc_lds_layout(acc_lds, S, SC_TERMS);
  acc_lds_zero(A, S, SC_TERMS);  // ends with a barrier: the shared cuts are visible
  const double tail = sh_st.tail, width = sh_st.width;
  const double inv_width = 1.0 / width;
  const double bx = sh_st.absmax[0], by = sh_st.absmax[2];
  const int ex = fx_exponent(bx), ey = fx_exponent(by);
  const int exx = fx_exponent(bx * bx), eyy = fx_exponent(by * by);
  constexpr int CH = EN_HIST_THREADS * EN_HIST_PPT;
  const int64_t nc = (n + CH - 1) / CH;
  double zz[EN_HIST_PPT], xx[EN_HIST_PPT], yy[EN_HIST_PPT];
  double zn[EN_HIST_PPT], xn[EN_HIST_PPT], yn[EN_HIST_PPT];
  auto load_chunk = [&](int64_t c, double (&zv)[EN_HIST_PPT], double (&xv)[EN_HIST_PPT],
                        double (&yv)[EN_HIST_PPT]) {
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      const int64_t j = c * CH + threadIdx.x + (int64_t)k * EN_HIST_THREADS;
      if (c < nc && j < n) {
        zv[k] = e.coords[4][j];
        xv[k] = e.coords[0][j];
        yv[k] = e.coords[2][j];
      } else {
        zv[k] = 0.0;
        xv[k] = 0.0;
        yv[k] = 0.0;
      }
    }
  };
  load_chunk(bid, zz, xx, yy);
  for (int64_t c = bid; c < nc; c += nblk) {
    load_chunk(c + nblk, zn, xn, yn);
    const int64_t base = c * CH + threadIdx.x;
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      const int64_t j = base + (int64_t)k * EN_HIST_THREADS;
      if (j < n) {
        const double f = slice_floor_r(zz[k], tail, width, inv_width);
        if (f >= 0.0 && f < (double)S) {
          const int s = (int)f;
          atomicAdd(&A.counts[s], 1u);
          acc_add_value(A, S, 0, s, xx[k], ex, P.fallback);
          acc_add_value(A, S, 1, s, yy[k], ey, P.fallback);
          acc_add_value(A, S, 2, s, xx[k] * xx[k], exx, P.fallback);
          acc_add_value(A, S, 3, s, yy[k] * yy[k], eyy, P.fallback);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < EN_HIST_PPT; ++k) {
      zz[k] = zn[k];
      xx[k] = xn[k];
      yy[k] = yn[k];
    }
  }
  acc_lds_store(A, S, SC_TERMS, P, bid);
}

// A fixed-point total as an unevaluated double-double hi + lo (exact):
// the integer h*2^32 + l, h = hi_sum + (lo_sum >> 32), l = lo_sum mod 2^32,
// times 2^-e.
__device__ __forceinline__ void fx_value_dd(unsigned long long hi_sum, unsigned long long lo_sum,
                                            int e, double &hi, double &lo) {
  const long long h = (long long)hi_sum + (long long)(lo_sum >> 32);
  const double a = (double)h;
  const long long r = h - (long long)a;  // |r| <= 2^10
  const double A = ldexp(a, 32);
  const double B = (double)(r * 4294967296LL + (long long)(lo_sum & 0xffffffffULL));
  const double s = A + B;
  const double bb = s - A;
  const double t = (A - (s - bb)) + (B - bb);
  hi = ldexp(s, -e);
  lo = ldexp(t, -e);
}

// (q - x^2/n) / (n - 1) from double-double totals x = sum u, q = sum u^2:
// the variance with the single rounding of each u^2 and of the result.
__device__ __forceinline__ double var_from_sums(double xh, double xl, double qh, double ql,
                                                double n) {
  // x^2 as double-double
  const double ph = xh * xh;
  const double pl = fma(xh, xh, -ph) + 2.0 * xh * xl;
  // (ph + pl) / n
  const double q1 = ph / n;
  const double r = fma(-q1, n, ph);
  const double q2 = (r + pl) / n;
  // (qh + ql) - (q1 + q2)
  const double s = qh - q1;
  const double bb = s - qh;
  const double err = (qh - (s - bb)) + (-q1 - bb);
  const double d = s + ((err + ql) - q2);
  const double v = d / (n - 1.0);
  return v > 0.0 ? v : 0.0;
}

// Merge of the space-charge partial histograms: per slice the count, means
// (mean_per_slice: sum / n, 0 for empty slices), standard deviations
// (std_per_slice: 0 for n <= 1) and the line density, into the field rows
// the next stage's kick reads.
template <int SL>
__device__ __forceinline__ void sc_merge_body(const EngineArgs &a, int stage, int bid) {
  const phx_engine &e = a.e;
  const ScWs sw = sc_ws(e.sc_ws, e.sc_slices);
  const int turn = *e.turn_counter;
  const phx_turn_program &g = e.programs[e.turn_program[turn] + stage];
  const int S = g.sc_slicer.n_slices;
  const AccParts P = acc_parts_layout(sw.acc, hist_blocks(e.n), S, SC_TERMS);
  SliceTotals tot;
  if (!acc_merge<SC_TERMS, SL>(P, S, tot, bid)) return;
  const int s = bid * SL + threadIdx.x % SL;
  const phx_slice_state &st = *sw.state;
  const double tail = st.tail, width = st.width;
  const double bx = st.absmax[0], by = st.absmax[2];
  const int ex = fx_exponent(bx), ey = fx_exponent(by);
  const int exx = fx_exponent(bx * bx), eyy = fx_exponent(by * by);
  const double cnt = (double)tot.count;
  const double fx = P.fallback[s], fy = P.fallback[S + s];
  const double fxx = P.fallback[2 * S + s], fyy = P.fallback[3 * S + s];
  P.fallback[s] = 0.0;
  P.fallback[S + s] = 0.0;
  P.fallback[2 * S + s] = 0.0;
  P.fallback[3 * S + s] = 0.0;
  const double sx = fx_value(tot.hi[0], tot.lo[0], ex) + fx;
  const double sy = fx_value(tot.hi[1], tot.lo[1], ey) + fy;
  ScSlice row;
  row.mx = tot.count ? sx / cnt : 0.0;
  row.my = tot.count ? sy / cnt : 0.0;
  double sgx = 0.0, sgy = 0.0;
  if (tot.count > 1) {
    double xh, xl, yh, yl, qxh, qxl, qyh, qyl;
    fx_value_dd(tot.hi[0], tot.lo[0], ex, xh, xl);
    fx_value_dd(tot.hi[1], tot.lo[1], ey, yh, yl);
    fx_value_dd(tot.hi[2], tot.lo[2], exx, qxh, qxl);
    fx_value_dd(tot.hi[3], tot.lo[3], eyy, qyh, qyl);
    sgx = sqrt(var_from_sums(xh, xl + fx, qxh, qxl + fxx, cnt));
    sgy = sqrt(var_from_sums(yh, yl + fy, qyh, qyl + fyy, cnt));
  }
  row.nl = sc_line_density((int32_t)tot.count, g.sc.charge_per_mp, arange_value(tail, width, s),
                           arange_value(tail, width, s + 1));
  row.pad = 0.0;
  row.f = sc_field_params(sgx, sgy, g.sc);
  sw.slices[s] = row;
}

// Slices per merge block: 4 up to 512 slices, 8 above (5.2 / 6.3 us at 500
// slices, 7.2 / 6.3 us at 1000); integer sums, the same results.
__host__ __device__ inline int merge_sl(int S) { return S <= 512 ? 4 : 8; }

// Wake convolution block (tile, ki, t) of the general kernel.
__device__ __forceinline__ void conv_body(const EngineArgs &a, int tile, int ki, int t) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const phx_turn_program &g = e.programs[e.turn_program[e.state->turn]];
  const WakeKicks &kk = *reinterpret_cast<const WakeKicks *>(g.kick);
  if (tile == 0 && ki == 0 && t == 0 && threadIdx.x == 0) *e.pending = 1;
  wake_conv_block(e.hist, kk, e.n_slices, w.conv_part, a.T, e.kicks, a.gmax, tile, ki, t,
                  a.n_kicks);
}

template <int KM>
__device__ __forceinline__ void conv_rw_body(const EngineArgs &a, int bid) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  if (bid == 0 && threadIdx.x == 0) *e.pending = 1;
  wake_rw_block<KM>(e.hist, engine_rw_desc(w, a.n_kicks, e.n_slices, a.T), a.n_kicks,
                    e.n_slices, w.conv_part, a.T, e.kicks, bid);
}

__device__ __forceinline__ void conv_reduce_body(const EngineArgs &a, int bid) {
  const phx_engine &e = a.e;
  EngineWs w = engine_ws(e.ws, e.n, e.n_slices);
  const phx_turn_program &g = e.programs[e.turn_program[e.state->turn]];
  const WakeKicks &kk = *reinterpret_cast<const WakeKicks *>(g.kick);
  wake_conv_reduce(e.hist, kk, a.n_kicks, e.n_slices, w.conv_part, a.T, e.kicks, bid);
}

// Launch arguments of engine e (checked): the record, its kick count and
// memory depth, the pass grid and the convolution blocks (phx_engine.hip).
int engine_args_for(const phx_engine *e, EngineArgs *out);

}  // namespace phx_en
