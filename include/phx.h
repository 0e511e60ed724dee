/*
 * phx.h - C ABI of the MI355X (gfx950) macroparticle one-turn-map library.
 *
 * The library replaces the per-turn hot path of PyHEADTAIL 1.16.5: the
 * transverse segment map with detuning, the longitudinal maps, uniform-bin
 * slicing with per-slice moments, and the wake-field convolution and kick.
 * It is the analogue of the reference's ctypes boundary to its Thrust
 * library (PyHEADTAIL/gpu/thrust_interface.py:5-269 <-> gpu/thrust_code.cu),
 * redesigned:
 *   - every function returns int (PHX_OK = 0, a negative PHX_E_* code, or a
 *     positive hipError_t value); phx_last_error() gives a message;
 *   - every buffer is owned by the caller (device memory), every call is
 *     asynchronous on the given stream (a hipStream_t passed as void*; NULL is
 *     the default stream), and no call allocates device memory except
 *     phx_graph_* (graph objects) and phx_init;
 *   - counts are int64_t (no 2^31 limit on particle numbers);
 *   - parameter records are plain C structs passed by host pointer and copied
 *     into the launch, so the caller may reuse them immediately.
 *
 * Coordinates are the six FP64 structure-of-arrays columns of a Particles
 * object (x, xp, y, yp, z, dp; particles/particles.py:16-229).
 */
#ifndef PHX_H
#define PHX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHX_ABI_VERSION 2

enum {
    PHX_OK = 0,
    PHX_E_INVALID = -1,   /* bad argument (null pointer, bad size, bad enum) */
    PHX_E_LIMIT = -2,     /* a compile-time limit below was exceeded */
    PHX_E_NODEVICE = -3,  /* no HIP device or phx_init not called */
    PHX_E_STATE = -4      /* inconsistent device state (e.g. empty wake history) */
};

#define PHX_MAX_DETUNERS 4       /* detuners per transverse segment */
#define PHX_MAX_CHROMA 8         /* chromaticity polynomial order */
#define PHX_MAX_LONG_OPS 10      /* drifts + kicks in one longitudinal sequence */
#define PHX_MAX_KICKS 16         /* wake kicks in one call / turn program */
#define PHX_MAX_TERMS 8          /* per-slice accumulated terms in one pass */
#define PHX_MAX_PROGRAM 16       /* local elements fused into one particle pass */
#define PHX_MAX_STATS 16         /* per-turn bunch statistics recorded by a pass */
#define PHX_STANDALONE_GRID 768  /* blocks of the reducing particle passes of the
                                    standalone calls (phx_bunch_sums,
                                    phx_slice_cuts): their fold order */
#define PHX_MAX_GRID 2048        /* largest engine pass grid */
#define PHX_ENGINE_MAX_N 536870911LL  /* macroparticles of one bunch in the fused
                                         engine (coordinates addressed by 32-bit
                                         byte offsets) */

typedef void *phx_stream_t;

/* ------------------------------------------------------------------------ */
/* Library state                                                            */
/* ------------------------------------------------------------------------ */

/* Message of the last failing call on this host thread. */
const char *phx_last_error(void);
/* ABI version (PHX_ABI_VERSION). */
int phx_abi_version(void);
/* Make `device` current for the calling thread and query its properties.  */
int phx_init(int device);
/* sizeof of a public record by name ("phx_tseg", "phx_turn_program", ...),
   -1 if unknown: lets bindings check their struct layouts. */
int64_t phx_sizeof(const char *type_name);
/* Macroparticles per group of the fused engine's particle pass: a pass of
   G blocks gives block b the groups b, b + G, ... (the last one partial),
   so the pass grid that gives every block k groups is
   ceil(ceil(n / group) / k).  No reference counterpart. */
int32_t phx_pass_group_size(void);
/* Bytes of workspace that the reduction and slicing calls need for n
   particles and up to n_slices slices.  The workspace must be zero-filled
   once when allocated; every call leaves it zero-filled again. */
int64_t phx_workspace_bytes(int64_t n, int32_t n_slices);

/* ------------------------------------------------------------------------ */
/* Bunch statistics  (Particles.mean_* / sigma_* / epsn_*:                  */
/* particles/particles.py:311-366; general/pmath.py:183-185;                */
/* cobra_functions/stats.pyx:17-122)                                        */
/* ------------------------------------------------------------------------ */

/*
 * Reduces k <= 6 device arrays a_0..a_{k-1} of length n in one pass and
 * writes to the device array `out` (length phx_bunch_sums_len(k)):
 *   out[0      .. k)      sum a_i                       (for np.mean)
 *   out[k      .. 2k)     sum (a_i - a_i[0])            (shifted sums of cov)
 *   out[2k     .. 2k+P)   sum (a_i - a_i[0])(a_j - a_j[0]), i <= j, row-major
 *   out[2k+P   .. 3k+P)   min a_i
 *   out[3k+P   .. 4k+P)   max a_i
 *   out[4k+P   .. 5k+P)   max |a_i|
 * with P = k(k+1)/2.  The reduction order is fixed (chunks of PHX chunk size,
 * then chunk order), so the result is bit-reproducible for a given n.
 * Replaces np.mean, cobra_functions/stats.pyx:cov (shift = first element)
 * and np.min/np.max used by Slicer.get_long_cuts (particles/slicing.py:395).
 */
int64_t phx_bunch_sums_len(int32_t k);
int phx_bunch_sums(const double *const *arrays, int32_t k, int64_t n,
                   double *out, void *ws, phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Uniform-bin slicing (particles/slicing.py:532-615, 58-317;              */
/* cobra_functions/stats.pyx:192-425)                                      */
/* ------------------------------------------------------------------------ */

enum { PHX_CUTS_FIXED = 0, PHX_CUTS_NSIGMA = 1, PHX_CUTS_MINMAX = 2 };

typedef struct {
    int32_t n_slices;
    int32_t cuts_mode;        /* PHX_CUTS_* */
    double n_sigma_z;         /* PHX_CUTS_NSIGMA */
    double z_cut_tail;        /* PHX_CUTS_FIXED */
    double z_cut_head;        /* PHX_CUTS_FIXED */
} phx_slicer;

/* Device-resident result of the cut computation of one slicing. */
typedef struct {
    double tail;              /* z_cut_tail */
    double head;              /* z_cut_head */
    double width;             /* (head - tail) / n_slices */
    double mean_z;            /* np.mean(z)   (PHX_CUTS_NSIGMA only) */
    double sigma_z;           /* cov(z, z)^.5 (PHX_CUTS_NSIGMA only) */
    double absmax[6];         /* max |coordinate| of x, xp, y, yp, z, dp over
                                 the finite values (the fused engine keeps
                                 those of x and y only; the others -inf) */
    int64_t n;                /* particles in the bunch */
    int32_t n_slices;
    int32_t turn;             /* turn index of the fused engine */
    int32_t valid;
    int32_t pad;
} phx_slice_state;

/*
 * Computes the slicing window of UniformBinSlicer.compute_sliceset_kwargs
 * (particles/slicing.py:599-615 with get_long_cuts :395-413) into the
 * device record *state, and z_bins (n_slices + 1 FP64, device) restating
 * np.arange(tail, head + 1e-7*width, width) (general/pmath.py:213-216).
 * Also records max|u| of the six coordinates (null pointers give 0), which
 * fixes the fixed-point scales of phx_slice_accumulate.
 */
int phx_slice_cuts(const double *const coords[6], int64_t n,
                   const phx_slicer *cfg, phx_slice_state *state,
                   double *z_bins, void *ws, phx_stream_t stream);

/* One accumulated per-slice term: sum over particles of slice s of
 *    (u_a - c_a[s]) * (u_b - c_b[s])      (b < 0: the single factor u_a - c_a[s])
 * where c_a / c_b are per-slice centres (null: 0).  coord indices 0..5 are
 * x, xp, y, yp, z, dp. */
typedef struct {
    int32_t a, b;
    const double *center_a;
    const double *center_b;
} phx_term;

/*
 * One pass over the particles of the slicing described by *state:
 *   idx = int32(floor((z - tail) / width))   (particles/slicing.py:609-611)
 * or, when idx_in is not null, the stored slice_index_of_particle (as the
 * reference's per-slice statistics use, stats.pyx:235-425); idx is
 * written to idx_out when not null (x86 conversion semantics: NaN and
 * out-of-range give INT32_MIN), counts[s] = number of particles with
 * 0 <= idx < n_slices (stats.pyx:192-202), and for each term t the exact
 * per-slice sum into sums[t*n_slices + s].  Sums use split 64-bit
 * fixed-point accumulators scaled from max|u|, so they are independent of
 * particle order and thread schedule (bit-reproducible).  When means is not
 * null, means[t*n_slices+s] = sums/counts for counts > 0, else 0
 * (mean_per_slice, stats.pyx:235-253).
 */
int phx_slice_accumulate(const double *const coords[6], int64_t n,
                         const phx_slice_state *state, int32_t n_slices,
                         const phx_term *terms, int32_t n_terms,
                         const int32_t *idx_in, int32_t *idx_out,
                         int32_t *counts,
                         double *sums, double *means,
                         void *ws, phx_stream_t stream);

/* Per-slice statistics from centred second-moment sums (device arrays of
 * length n_slices):  kind 0: std  = sqrt(c_aa/(n-1))        (stats.pyx:258-286)
 *                    kind 1: cov  = c_ab/(n-1)              (stats.pyx:290-319)
 *                    kind 2: emittance from 3 (dp==null) or 6 sums
 *                            (stats.pyx:374-425)
 * sums holds the needed centred sums in the order
 * (uu, uup, upup[, udp, updp, dpdp]).  `factor` multiplies the result
 * The result is ((f_pre * r) * f_mul) / f_div: (1, betagamma, 1) for
 * epsn_x/y and (4*pi, p0, e) for epsn_z (particles/slicing.py:502-518). */
int phx_slice_finalize(int32_t kind, int32_t n_slices, const int32_t *counts,
                       const double *sums, int32_t n_sums, double f_pre,
                       double f_mul, double f_div, double *out,
                       phx_stream_t stream);

/* z_centers and times of a slicing: z_centers = z_bins[:-1] + .5*diff(z_bins)
 * (particles/slicing.py:124-126), times = z_centers / (beta*c) with
 * beta_c = beta * c evaluated by the caller (:303-307). */
int phx_slice_times(const double *z_bins, int32_t n_slices, double beta_c,
                    double *z_centers, double *times, phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Transverse segment map (trackers/transverse_tracking.py:17-196,         */
/* trackers/detuners.py:48-151)                                            */
/* ------------------------------------------------------------------------ */

enum { PHX_DET_CHROMA = 1, PHX_DET_AMPLITUDE = 2 };

typedef struct {
    int32_t type;             /* PHX_DET_* */
    int32_t order;            /* chromaticity: number of coefficients */
    /* chromaticity: c_i = Qp_i * dmu / (i+1)!  (detuners.py:83) */
    double cx[PHX_MAX_CHROMA];
    double cy[PHX_MAX_CHROMA];
    /* amplitude detuning (detuners.py:139-151); host-evaluated factors
       gx = (1 + alpha_x^2)/beta_x, ax2 = 2*alpha_x, bx = beta_x, same in y */
    double gx, ax2, bx, gy, ay2, by;
    double dapp_x, dapp_y, dapp_xy, dapp_yx;
    double p0;                /* beam.p0 at track time */
    double inv_p0;            /* 1.0 / p0 (IEEE division) */
} phx_detuner;

typedef struct {
    /* I and J of _build_segment_map (transverse_tracking.py:71-103) */
    double I00, I01, I10, I11, I22, I23, I32, I33;
    double J00, J01, J10, J11, J22, J23, J32, J33;
    /* M = I cos + J sin evaluated on the host when there is no detuner
       (transverse_tracking.py:178-193) */
    double M00, M01, M10, M11, M22, M23, M32, M33;
    double dQ_x, dQ_y;
    double D_x_s0, D_x_s1, D_y_s0, D_y_s1;
    int32_t with_dispersion;  /* _track_with_dispersion (:105-125) */
    int32_t n_detuners;
    phx_detuner det[PHX_MAX_DETUNERS];
} phx_tseg;

/* TransverseSegmentMap.track (transverse_tracking.py:138-196). */
int phx_transverse_segment(double *x, double *xp, double *y, double *yp,
                           const double *dp, int64_t n, const phx_tseg *p,
                           phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Longitudinal maps (trackers/longitudinal_tracking.py)                   */
/* ------------------------------------------------------------------------ */

enum { PHX_LONG_DRIFT = 1, PHX_LONG_KICK = 2 };

typedef struct {
    int32_t type;             /* PHX_LONG_* */
    int32_t with_dispersion;  /* Kick.track_with_dispersion (:185-193) */
    /* drift: z <- beta_ratio*z - (eta0*dp)*length       (:107-111, :62-81) */
    double beta_ratio, eta0, length;
    /* kick: phi = (h*((2pi*z)/C) + phi_offset) + phi_lock
             dp <- (dp*p0_before + ((amplitude*sin(phi)) - p_increment))
                   / p0_after                               (:195-203) */
    double harmonic, two_pi, circumference, phi_offset, phi_lock;
    double amplitude, p_increment, p0_before, p0_after;
    double D_x, D_y;
    /* 1.0 / circumference and 1.0 / p0_after (IEEE divisions): the kick's
       two divisions are formed from them and rounded exactly as the
       divisions themselves */
    double inv_circumference, inv_p0_after;
} phx_long_op;

typedef struct {
    int32_t n_ops;
    int32_t shrink;           /* RFSystems._shrink_transverse_emittance */
    double shrink_factor;     /* sqrt(betagamma_old/betagamma) (:556-565) */
    phx_long_op op[PHX_MAX_LONG_OPS];
} phx_long_seq;

/* Drift.track, Kick.track and RFSystems.track (drift, kicks, drift;
   longitudinal_tracking.py:84-203, 362-369, 556-574) as one sequence. */
int phx_longitudinal(double *x, double *xp, double *y, double *yp,
                     double *z, double *dp, int64_t n, const phx_long_seq *p,
                     phx_stream_t stream);

typedef struct {
    double cosQ, sinQ, longfac;   /* host-evaluated as in :743-757 */
    double D_x, D_y;
} phx_linmap;

/* LinearMap.track -> track_with_dispersion (longitudinal_tracking.py:
   720-759). */
int phx_linear_map(double *x, double *y, double *z, double *dp, int64_t n,
                   const phx_linmap *p, phx_stream_t stream);

/* Ideal transverse damper (feedback/transverse_damper.py:15-94): with
   m_u = sums[u] / n the bunch means from phx_bunch_sums over (xp, yp, x, y)
   (sums: device, first four entries),
     xp -= coef_xp * m_xp;   yp -= coef_yp * m_yp            (resistive part)
     xp -= (coef_x * m_x) / beta;   yp -= (coef_y * m_y) / beta  (if beta)
   in that order, per active plane.  The coefficients are the host products
   gain * sin(phase) and gain * cos(phase), gain = 2 / damping rate. */
typedef struct {
    double coef_xp, coef_yp;  /* gain * sin(phase) per plane */
    double coef_x, coef_y;    /* gain * cos(phase) per plane */
    double beta;              /* local_beta_function; 0: no reactive part */
    int32_t planes;           /* bit 0: horizontal, bit 1: vertical */
    int32_t pad;
} phx_damper;
int phx_transverse_damper(double *xp, double *yp, int64_t n, const double *sums,
                          const phx_damper *p, phx_stream_t stream);

/* The FP64 sine and cosine the maps evaluate on the device (the detuned
   betatron phase, the RF phase): s[i] = sin(x[i]), c[i] = cos(x[i]) for
   device arrays of n values; either output may be null.  No reference
   counterpart: it exposes the device arithmetic behind
   TransverseSegmentMap.track and Kick.track (np.sin / np.cos,
   general/pmath.py:178-179, 239) for accuracy checks. */
int phx_math_sincos(const double *x, int64_t n, double *s, double *c,
                    phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Wake fields (impedances/wakes.py:77-145, 379-678;                        */
/* impedances/wake_kicks.py:19-361)                                         */
/* ------------------------------------------------------------------------ */

enum {
    PHX_WAKE_RES_TRANSVERSE = 1,  /* Resonator.function_transverse (:450-470) */
    PHX_WAKE_RES_LONGITUDINAL = 2,/* Resonator.function_longitudinal (:472-493) */
    PHX_WAKE_RESISTIVE_WALL = 3,  /* ResistiveWall.function_transverse (:611-624) */
    PHX_WAKE_TABLE = 4            /* WakeTable.function_transverse /
                                     function_longitudinal (:290-376) */
};
/* Wake tables: linear interpolation of (table_t, table_w) at -dt with
   np.interp's arithmetic (interp1d delegates to it for FP64 data).
   TRANS_UR: ultrarelativistic transverse, W(-min(dt, 0)); TRANS_LOWBETA:
   W(-dt); both NaN outside the table (the reference raises ValueError).
   LONG_BL: longitudinal with table_t[0] == 0, (sign(-dt) + 1)/2 * W(-dt);
   LONG: W(-dt); both 0 outside the table.  table_w holds the strengths
   already converted to SI units with the reference's signs. */
enum { PHX_TABLE_TRANS_UR = 0, PHX_TABLE_TRANS_LOWBETA = 1, PHX_TABLE_LONG_BL = 2,
       PHX_TABLE_LONG = 3 };
enum { PHX_MOM_N = 0, PHX_MOM_X = 1, PHX_MOM_Y = 2 };
enum { PHX_PLANE_XP = 0, PHX_PLANE_YP = 1, PHX_PLANE_DP = 2 };
/* Arithmetic of a multi-turn (dot-product) resistive-wall convolution:
     FP64         FP64; the source turns gone by (every pair beyond |dt_min|,
                  time offsets within 1e-3 of the distance of the two
                  slicings) through a six-term separable expansion of
                  (-dt)^(-1/2), exact to below 1e-18 of the sum of the
                  terms' magnitudes (O(n_slices) per turn); the current turn
                  pair by pair;
     FP32         FP32 wake values (from the FP64 time differences) and FP32
                  products summed in FP32 over each run of at most
                  ceil(n_slices/8) sources, those partial sums added in FP64;
     FP64_DIRECT  FP64 pair by pair for every turn (the reference's sums);
     MFMA_F32     every pair on the matrix cores: 4x4x1 FP32 MFMA, FP32 wake
                  values and moments, FP32 accumulation over 256 sources,
                  FP64 beyond;
     MFMA_BF16X3  the same with 4x4x4 bfloat16 MFMA on split operands: two
                  bfloat16 parts of each wake value and moment, three
                  products (about 16 significant bits);
     MFMA_BF16X6  three parts each, six products (two MFMAs per step).
   Every other convolution is FP64 whatever the field says. */
enum { PHX_PREC_FP64 = 0, PHX_PREC_FP32 = 1, PHX_PREC_FP64_DIRECT = 2, PHX_PREC_MFMA_F32 = 3,
       PHX_PREC_MFMA_BF16X3 = 4, PHX_PREC_MFMA_BF16X6 = 5 };

typedef struct {
    int32_t fn;               /* PHX_WAKE_* */
    int32_t moment;           /* PHX_MOM_*: source moment n, n*mean_x, n*mean_y */
    int32_t plane;            /* PHX_PLANE_*: momentum that receives the kick */
    int32_t multiply;         /* -1: += k[idx];  0: += k[idx]*x;  2: += k[idx]*y */
    int32_t n_turns_wake;
    int32_t dot_product;      /* 1: _convolution_dot_product, 0: np.convolve */
    int32_t precision;        /* PHX_PREC_* */
    int32_t reserved;         /* 0 */
    /* resonator: branch by Q (>0.5, ==0.5, <0.5); prefix = host-evaluated
       Yokoya*R*omega^2/(Q*omegabar) (transverse) or R (longitudinal) */
    double Q, prefix, alpha, omegabar, alpha_over_omegabar;
    /* resistive wall */
    double yokoya, L, b3, dt_min_abs, k_sigma, sqrt_z0c;
    double wake_factor;       /* WakeKick._wake_factor of the current bunch */
    /* wake table (PHX_WAKE_TABLE): device arrays of table_n >= 2 points,
       table_t strictly increasing */
    const double *table_t, *table_w;
    int32_t table_n;
    int32_t table_kind;       /* PHX_TABLE_* */
} phx_wake_kick;

/* Multi-turn history of slice data (impedances/wakes.py:116-145): a ring of
   `capacity` slots; slot (head - i) mod capacity is the SliceSet i turns old.
   All pointers are device memory.  The record itself lives in device memory
   so that the fused engine can advance it without host round trips. */
typedef struct {
    int32_t n_slices;
    int32_t capacity;
    int32_t n_filled;
    int32_t head;
    double *times;            /* [capacity][n_slices] z_centers/(beta c) */
    double *mom;              /* [capacity][3][n_slices] n, n*mean_x, n*mean_y */
    double *age;              /* [capacity] */
    double *beta;             /* [capacity] SliceSet.beta of each entry */
} phx_wake_hist;

/* Pushes the slice data of a new SliceSet into the device history *hist:
   ages += age_increment (C/(beta c) of the current bunch, wakes.py:135-137),
   then appendleft (times, counts, mean_x, mean_y) with age 0 (:141-142). */
int phx_wake_hist_push(phx_wake_hist *hist, const double *times,
                       const int32_t *counts, const double *mean_x,
                       const double *mean_y, double age_increment,
                       double beta, phx_stream_t stream);

/* Wake signals of n_kicks kicks into kicks_out[k*n_slices + j]:
   wake_factor * sum over the stored turns of the convolution of the source
   moments with the wake function (wake_kicks.py:75-139), in FP64 with the
   reference's arithmetic for the wake functions.  Kick k uses
   min(n_filled, n_turns_wake, max_turns) stored turns; dot_product selects
   _convolution_dot_product (dt[i,j] = t_tgt[j] - (t_src[i] + age)) instead
   of the np.convolve 'valid' form.  ws: phx_wake_workspace_bytes bytes,
   zero-filled once (calls leave it zero-filled). */
int64_t phx_wake_workspace_bytes(int32_t n_kicks, int32_t max_turns,
                                 int32_t n_slices);
int phx_wake_convolve(const phx_wake_hist *hist, const phx_wake_kick *kicks,
                      int32_t n_kicks, int32_t n_slices, int32_t max_turns,
                      double *kicks_out, void *ws, phx_stream_t stream);

/* Applies the kicks to the particles inside the cuts, in list order:
   p[idx] += k[idx] (* x or * y)   (wake_kicks.py:160-361).  The slice index
   comes from idx when not null, else it is recomputed from z and *state
   (bit-identical). */
int phx_wake_apply(double *const coords[6], int64_t n, const int32_t *idx,
                   const phx_slice_state *state, const phx_wake_kick *kicks,
                   int32_t n_kicks, int32_t n_slices, const double *kicks_in,
                   phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Uniform-charge slicing (particles/slicing.py:618-686)                    */
/* ------------------------------------------------------------------------ */

/* Workspace of phx_ucs_sort for n particles (zero-filling not needed). */
int64_t phx_ucs_workspace_bytes(int64_t n);
/* z sorted ascending (z_sorted) with the permutation perm (z_sorted[r] =
   z[perm[r]]; stable: equal z keep their index order, where np.argsort's
   quicksort leaves it unspecified), and below[0] = #(z < tail), below[1] =
   #(z < head) for the cuts in *state (np.searchsorted on the sorted z).
   below: device, 2 x uint64. */
int phx_ucs_sort(const double *z, int64_t n, const phx_slice_state *state, double *z_sorted,
                 int32_t *perm, uint64_t *below, void *ws, phx_stream_t stream);
/* Slice indices and bin edges from first (device, n_slices + 1 entries):
   sorted position r lies in slice i for first[i] <= r < first[i+1], else
   -1; idx[perm[r]] is written for every r.  z_bins[0] and z_bins[n_slices]
   are the cuts, z_bins[i] = (z_sorted[first[i]-1] + z_sorted[first[i]])/2
   (NaN where first[i] is not a particle position). */
int phx_ucs_assign(const double *z_sorted, const int32_t *perm, int64_t n, int32_t n_slices,
                   const int32_t *first, const phx_slice_state *state, int32_t *idx,
                   double *z_bins, phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* 2.5-D transverse Gaussian space charge (spacecharge/spacecharge.py:      */
/* 159-417)                                                                 */
/* ------------------------------------------------------------------------ */

/* The slice data TransverseGaussianSpaceCharge.track reads (device): the
   counts and bin edges (Nlambda = counts * charge_per_mp / diff(z_bins)),
   the slice means and standard deviations of x and y. */
typedef struct {
    int32_t n_slices;
    int32_t pad;
    const int32_t *counts;    /* [n_slices] */
    const double *z_bins;     /* [n_slices + 1] */
    const double *mean_x, *mean_y, *sigma_x, *sigma_y;  /* [n_slices] */
} phx_sc_slices;

typedef struct {
    double prefactor;         /* charge * length / (p0 * betagamma * gamma * c) */
    double charge_per_mp;
    double denom_pre;         /* 2 * epsilon_0 * sqrt(pi) */
    double round_pre;         /* 2 * pi * epsilon_0 */
    double rtol, atol;        /* ratio_threshold, absolute_threshold */
    int32_t sig_check;        /* 1: swap x/y for sigma_x < sigma_y, round beam */
    int32_t pad;
} phx_spacecharge;

/* TransverseGaussianSpaceCharge.track on the particles inside the cuts:
   xp += (E_x/Q * Nlambda) * prefactor, yp likewise, with the Bassetti-
   Erskine field of the particle's slice (Faddeeva function by Weideman's
   rational approximation, N = 40).  The slice index comes from idx when not
   null, else from z and *state. */
int phx_spacecharge_kick(double *const coords[6], int64_t n, const int32_t *idx,
                         const phx_slice_state *state, const phx_sc_slices *slices,
                         const phx_spacecharge *p, phx_stream_t stream);

/* The device Faddeeva function: wr + i wi = w(x + i y) for y >= 0 (device
   arrays of n values).  No reference counterpart: it exposes the function
   behind the space-charge field (pmath.wofz -> scipy.special.wofz,
   general/pmath.py:50-52) for accuracy checks. */
int phx_math_wofz(const double *x, const double *y, int64_t n, double *wr, double *wi,
                  phx_stream_t stream);

/* ------------------------------------------------------------------------ */
/* Fused one-turn map                                                       */
/* ------------------------------------------------------------------------ */

enum { PHX_EL_TSEG = 1, PHX_EL_LONG = 2, PHX_EL_LINMAP = 3 };

/* Parameters of the local (per-particle) elements of one turn, in order. */
typedef struct {
    int32_t n_el;
    int32_t el_type[PHX_MAX_PROGRAM];
    int32_t el_index[PHX_MAX_PROGRAM];
    int32_t n_tseg, n_long, n_lin;
    int32_t has_wake;
    phx_tseg tseg[4];
    phx_long_seq lseq[2];
    phx_linmap lin[2];
    /* wake field ending the turn */
    phx_slicer slicer;
    double beta;              /* beam.beta at slicing time (SliceSet.beta) */
    double beta_c;            /* beam.beta * c at slicing time */
    double age_increment;     /* C/(beta c) at WakeField.track time */
    int32_t n_kicks;
    int32_t pad;
    phx_wake_kick kick[PHX_MAX_KICKS];
    /* space charge ending this stage (turns split into stages at their
       space-charge kicks, see phx_engine.sc_stages): has_sc = 1 when a
       TransverseGaussianSpaceCharge kick follows this stage's elements;
       sc_slicer its slicing, sc its parameters.  The kick itself is applied
       at the start of the next stage's particle pass. */
    int32_t has_sc;
    int32_t pad_sc;
    phx_slicer sc_slicer;
    phx_spacecharge sc;
} phx_turn_program;

/* Per-turn bunch statistics recorded at the end of each turn. */
enum {
    PHX_STAT_MEAN_X = 0, PHX_STAT_MEAN_XP, PHX_STAT_MEAN_Y, PHX_STAT_MEAN_YP,
    PHX_STAT_MEAN_Z, PHX_STAT_MEAN_DP, PHX_STAT_SIGMA_X, PHX_STAT_SIGMA_Y,
    PHX_STAT_SIGMA_Z, PHX_STAT_SIGMA_DP, PHX_STAT_EPSN_X, PHX_STAT_EPSN_Y,
    PHX_STAT_EPSN_Z, PHX_STAT_COUNT
};

/* Element orders of the turn programs that have a pass compiled for them
   (any other order: PHX_SHAPE_GENERIC, which runs every program).  RF1 / RF2
   is a longitudinal sequence [drift, 1 or 2 kicks, drift] (RFSystems with
   one or two harmonics), LIN a linear map, T one transverse segment. */
enum {
    PHX_SHAPE_GENERIC = 0,
    PHX_SHAPE_RF1_T = 1, PHX_SHAPE_RF2_T = 2, PHX_SHAPE_T_RF1 = 3, PHX_SHAPE_T_RF2 = 4,
    PHX_SHAPE_LIN_T = 5, PHX_SHAPE_T_LIN = 6, PHX_SHAPE_T = 7,
    PHX_SHAPE_NONE = 8,       /* no elements (a space-charge stage of the kick alone) */
    PHX_SHAPE_COUNT = 9
};
/* Flag added to PHX_SHAPE_RF1_T: every program's RF sequence is stationary,
   its drifts with beta_ratio = 1 and its kick with phi_offset = phi_lock =
   p_increment = +0 and no dispersion.  The pass then drops those terms,
   which leave the values unchanged (1 * z = z, (a + 0) + 0 = a + 0,
   b - 0 = b): the same bits as the general form. */
#define PHX_SHAPE_STATIONARY 16
/* Space-charge kicks per turn of the fused engine. */
#define PHX_MAX_SC_STAGES 16
/* Wake kick sets with a compiled kick step (other sets: PHX_KICKS_GENERIC):
   DXDY is [dipole x (xp += k), dipole y (yp += k)], as a round chamber's
   resonator or resistive wall gives; PLATES is [dipole x, quadrupole x
   (xp += k x), dipole y, quadrupole y (yp += k y)], a parallel-plates
   chamber's. */
enum { PHX_KICKS_GENERIC = 0, PHX_KICKS_DXDY = 1, PHX_KICKS_PLATES = 2, PHX_KICKS_COUNT = 3 };
/* Detuner sets of the transverse segments with compiled code (other sets:
   PHX_DETS_GENERIC): none, first-order chromaticity, first-order
   chromaticity then amplitude detuning (Synchrotron's order). */
enum { PHX_DETS_GENERIC = 0, PHX_DETS_NONE = 1, PHX_DETS_C1 = 2, PHX_DETS_C1_AMP = 3,
       PHX_DETS_COUNT = 4 };
/* Flag added to PHX_DETS_C1 / PHX_DETS_C1_AMP: every segment is smooth,
   I = identity, J00 = J11 = J22 = J33 = 0 exactly and no dispersion (equal
   optics at both ends with alpha = 0, as smooth-optics machines have), so
   M00 = I00 cos + J00 sin is cos bit for bit and the segment needs only
   J01, J10, J23, J32; an amplitude detuner then also has ax2 = ay2 = 0 and
   gx, gy > 0, and its 2 alpha u u' term is dropped (a zero added to
   gx u^2 >= +0). */
#define PHX_DETS_SMOOTH 8
/* Wake convolution kernels: resistive-wall kicks in the dot-product form
   (multi-turn memory) have their own, RW_DOT for the VALU precisions and
   RW_MFMA for the matrix-core ones; every other kick is GENERAL. */
enum { PHX_CONV_GENERAL = 1, PHX_CONV_RW_DOT = 2, PHX_CONV_RW_MFMA = 4 };

/* Everything the fused engine needs; all pointers are device memory owned
   by the caller.  The engine handles one-turn maps of the form
   [local elements..., WakeField] (or local elements only: n_slices = 0). */
typedef struct {
    double *coords[6];                  /* x, xp, y, yp, z, dp */
    int64_t n;
    int32_t n_slices;                   /* 0: no wake field */
    int32_t n_kicks;                    /* wake kicks of the wake field */
    int32_t hist_capacity;              /* deepest n_turns_wake (ring slots) */
    int32_t record_stats;               /* 1: fill stats[turn][33] */
    int32_t program_shape;              /* PHX_SHAPE_*: every program's element
                                           order (GENERIC is always valid) */
    int32_t kick_shape;                 /* PHX_KICKS_*: the wake field's kicks */
    int32_t pass_grid;                  /* blocks of the particle pass; 0: as
                                           many as stay resident on the device.
                                           The bunch sums are folded per block,
                                           so the last bits of the cuts depend
                                           on it: PHX_STANDALONE_GRID gives the
                                           cuts of phx_slice_cuts bit for bit */
    int32_t conv_kinds;                 /* PHX_CONV_* bits: the convolution
                                           kernels the programs' kicks need;
                                           0: both (always valid) */
    int32_t det_shape;                  /* PHX_DETS_*: the detuners of every
                                           program's transverse segments
                                           (GENERIC is always valid) */
    int32_t sc_stages;                  /* space-charge kicks per turn (0: none).
                                           With m > 0 every turn is m + 1
                                           stages, each a program of its own:
                                           turn_program[t] + k is stage k of
                                           turn t (programs of stages 0..m-1
                                           have has_sc = 1, stage m's 0);
                                           program_shape is then stage 0's
                                           element order, sc_stage_shape[k]
                                           stage k's, det_shape that of every
                                           stage's transverse segments */
    const phx_turn_program *programs;   /* [n_programs] */
    const int32_t *turn_program;        /* [n_turns_total] -> program index */
    double *stats;                      /* raw moment sums per turn, see below */
    int32_t *turn_counter;              /* device: index of the next turn */
    int32_t *pending;                   /* device flag: kicks not yet applied */
    phx_slice_state *state;             /* device: cuts of the last slicing */
    phx_wake_hist *hist;                /* device: multi-turn history */
    double *z_bins;                     /* [n_slices+1] */
    double *kicks;                      /* [n_kicks][n_slices] pending kick */
    int32_t *counts;                    /* [n_slices] of the last slicing */
    double *slice_means;                /* [2][n_slices] mean_x, mean_y */
    void *ws;                           /* phx_engine_ws_bytes, zeroed once */
    int32_t sc_slices;                  /* most slices of any space-charge
                                           slicing (0 without space charge) */
    int32_t pad_sc;
    void *sc_ws;                        /* phx_engine_sc_ws_bytes (space charge
                                           only), set up by phx_engine_reset */
    int32_t sc_stage_shape[PHX_MAX_SC_STAGES + 1];  /* PHX_SHAPE_* of stages
                                           1..sc_stages (entry 0 unused) */
    int32_t pad_sc2;
} phx_engine;

/* stats row r (state after r turns, r = 0 .. turns run) holds 33 FP64 sums
   over the bunch: [0,6) sum u for u = x, xp, y, yp, z, dp; [6,12)
   sum (u - u0); [12,33) sum (u - u0)(v - v0) for u <= v in that order, with
   u0 the coordinate of particle 0 (the shifts of stats.pyx:cov). */

/* Workspace bytes of the fused engine. */
int64_t phx_engine_ws_bytes(int64_t n, int32_t n_slices, int32_t n_kicks,
                            int32_t max_turns);
/* Space-charge workspace bytes of the fused engine: the space-charge cuts,
   the per-slice field parameters and the partial histograms of the
   space-charge slicing (sc_slices slices).  The space-charge stages of a
   turn (TransverseGaussianSpaceCharge.track, spacecharge/spacecharge.py:
   202-224, fused): stage k's particle pass applies the kick that ends stage
   k - 1, then the stage's elements; after it a slice histogram (counts and
   exact fixed-point sums of x, y, x^2, y^2 per slice on the cuts of the
   current z) and a merge (slice means, standard deviations
   sqrt((sum u^2 - (sum u)^2 / n) / (n - 1)) for n > 1, else 0, and the line
   density) give the field of the next kick. */
int64_t phx_engine_sc_ws_bytes(int64_t n, int32_t sc_slices);
/* Device-side reset of the engine state (turn counter, pending flag,
   workspace accumulators, wake history).  No memset nodes. */
int phx_engine_reset(const phx_engine *e, phx_stream_t stream);
/* Enqueue n_turns fused turns.  Each turn is four launches: the particle
   pass (pending wake kick, local elements, z moments), the slice histogram
   (cuts, per-block histograms), the histogram merge (slice means, history
   entry) and the wake convolution; without a wake field, the pass and a
   one-block finalize; one more launch ahead of the first turn records
   particle 0 (phx_engine_snapshot).  The wake kick of the last turn stays
   pending until phx_engine_flush. */
int phx_engine_turns(const phx_engine *e, int32_t n_turns, phx_stream_t stream);
/* Applies a pending wake kick (and records the final statistics row). */
int phx_engine_flush(const phx_engine *e, phx_stream_t stream);
/* Enqueue one stage of a turn: 0 particle pass, 1 slice histogram and
   merge, 2 wake convolution (stages 1-2 need a wake field).  Running 0, 1, 2
   in order is one turn; used to time the stages separately.  Call
   phx_engine_snapshot before the first stage 0 after the coordinates
   changed outside the engine. */
int phx_engine_stage(const phx_engine *e, int32_t stage, phx_stream_t stream);
/* Records particle 0's coordinates in the workspace: the shifts of the
   shifted sums of the next particle pass (stats.pyx:cov takes the first
   particle), read there by every block while block 0 rewrites particle 0.
   phx_engine_turns and phx_engine_flush call it themselves; graph replays
   do not. */
int phx_engine_snapshot(const phx_engine *e, phx_stream_t stream);

/* Allocation-free variant of phx_engine_turns.  It has no counterpart in
   the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  e->ws is ignored: the workspace is the library's own static
   device scratch of PHX_NOALLOC_SCRATCH_BYTES per device (held until the
   process ends); an engine whose phx_engine_ws_bytes exceeds it is
   rejected with PHX_E_LIMIT.  flags: PHX_NOALLOC_RESET resets the engine
   state first (as phx_engine_reset), PHX_NOALLOC_FLUSH applies the pending
   wake kick after the turns (as phx_engine_flush).  Every call on a device
   shares that device's scratch and leaves it in the state of its last
   turn: calls that overlap on one device (other streams, other threads)
   race.  Needs phx_init(device) first.  No hipMalloc, hipFree or memset. */
#define PHX_NOALLOC_SCRATCH_BYTES (24 << 20)
enum { PHX_NOALLOC_RESET = 1, PHX_NOALLOC_FLUSH = 2 };
int phx_engine_turns_noalloc(const phx_engine *e, int32_t n_turns, int32_t flags,
                             phx_stream_t stream);

/* Batches (the batched scan): n engines (independent bunches of any
   macroparticle count, slice count, wake memory and kick set; each with its
   own buffers, wake field required) advanced together, each stage of a turn
   one launch covering every engine (block offsets per engine in a device
   descriptor table), each engine's results those of its run alone.  The
   engines must share record_stats, program_shape and det_shape.  No
   counterpart in the reference, whose scans run one bunch after another
   (machines/synchrotron.py:294-298 per bunch).
   phx_batch_create copies the descriptor table (phx_batch_desc_bytes(n),
   caller-owned device memory) on `stream` and returns a host handle; the
   engines' buffers must stay where they are while it is used.
   phx_batch_turns: particle 0 of every engine, then n_turns turns.
   phx_batch_capture: a HIP graph of n_turns turns (replayed with
   phx_graph_launch, destroyed with phx_graph_destroy), without the
   particle-0 record: phx_batch_snapshot before the first replay after
   coordinates changed outside the engines.  The pending kick of each engine
   is applied by phx_engine_flush of that engine. */
int64_t phx_batch_desc_bytes(int32_t n);
int phx_batch_create(const phx_engine *engines, int32_t n, void *desc, int64_t desc_bytes,
                     phx_stream_t stream, void **batch);
int phx_batch_turns(void *batch, int32_t n_turns, phx_stream_t stream);
int phx_batch_snapshot(void *batch, phx_stream_t stream);
int phx_batch_capture(void *batch, int32_t n_turns, void **graph);
int phx_batch_destroy(void *batch);

/* HIP graph of n_turns fused turns (captured on an internal stream);
   replay with phx_graph_launch.  Returns an opaque handle.  The graph does
   not record particle 0 (each turn refreshes it for the next): call
   phx_engine_snapshot before the first replay after the coordinates changed
   outside the engine. */
int phx_graph_capture(const phx_engine *e, int32_t n_turns, void **graph);
int phx_graph_launch(void *graph, phx_stream_t stream);
int phx_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif

#endif /* PHX_H */
