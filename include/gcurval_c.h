/*
 * gcurval_c.h — C ABI of the MI355X (gfx950) arc-length engine.
 *
 * Plain pointers and sizes only. Every call returns a gcv_status (0 = OK);
 * gcv_last_error() gives the message of the last failure on a context. One
 * context per host thread; all work is stream-ordered on the context's HIP
 * stream, and results handed to the host synchronise that stream.
 *
 * Each entry point names the reference interface it replaces. Paths are
 * relative to the HolyChen/gcurval tree.
 */
#ifndef GCURVAL_C_H
#define GCURVAL_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gcv_status {
    GCV_OK = 0,
    GCV_ERR_INVALID = 1,     /* bad argument or batch layout */
    GCV_ERR_OOM = 2,         /* arena exhausted */
    GCV_ERR_HIP = 3,         /* HIP runtime error */
    GCV_ERR_UNSUPPORTED = 4, /* option combination not built */
    GCV_ERR_STATE = 5,       /* call order (e.g. sampling before a build) */
    GCV_ERR_DATA = 6         /* batch data: a non-finite knot, control point or semi-axis, or
                                knots that decrease within a curve (checked at table creation) */
} gcv_status;

/* Mirrors GcurvalEnum (gcurval/base/util.h:71-77). */
typedef enum gcv_where { GCV_DEVICE = 0, GCV_HOST = 1 } gcv_where;
typedef enum gcv_what { GCV_RETURN_US = 2, GCV_RETURN_POINTS = 3 } gcv_what;

typedef enum gcv_precision { GCV_FP32 = 0, GCV_FP64 = 1 } gcv_precision;

typedef struct gcv_ctx gcv_ctx;
typedef struct gcv_table gcv_table;

/*
 * Table-build options. Defaults (gcv_opts_default) follow gcurval/config.h:17-43
 * and gcurval/base/functors.inl.h:42: FP32, 3-point Gauss-Legendre, 64 root
 * sub-intervals per knot interval, depth 14 (FP32) / 23 (FP64), tolerance
 * 1e-6 (FP32) / 1e-14 (FP64), absolute tolerance on raw coordinates.
 */
typedef struct gcv_opts {
    int precision;     /* gcv_precision */
    int gl_order;      /* 3 or 5 */
    int n_subdivided;  /* roots per knot interval: power of two, 1..64 */
    int max_depth;     /* 0 = default; <= 18 (FP32), <= 31 (FP64) */
    double tolerance;  /* 0 = default; per-interval absolute tolerance */
    int normalize;     /* 1: test each curve at unit max-abs coordinate (lengths rescaled) */
    int newton_iters;  /* 0 = default: Newton steps after the Hermite start (FP32 mode: 1, or 2 without the FP64
                          step, then 1 FP64 step; FP64 mode: 3); lanes whose last step is still large go on */
    int skip_fp64_step; /* FP32 mode: 1 = no final FP64 Newton step (default 0: the step is taken) */
    int reserved[5];
} gcv_opts;

void gcv_opts_default(gcv_opts* o);

/* ---- context, arena, copies (replaces gMalloc/gFree/gFree_s over RMM or
 * cudaMalloc, gcurval/detail/cuda/util.h:41-71, and the cc()/cet() checks
 * that exit(-1), util.h:14-39) ----
 * arena_bytes: 0 = one hipMalloc per gcv_alloc (gcv_free synchronises the
 * stream and frees); GCV_ARENA_POOL = a pool that grows by hipMalloc chunks of
 * at least 256 MiB, kept until the context is destroyed (the RMM pool of the
 * reference's default build); any other value = one fixed chunk of that size.
 * In both arena forms gcv_free returns the block for reuse by later work on
 * the context's stream, without a synchronisation. */
#define GCV_ARENA_POOL ((size_t)-1)
int gcv_ctx_create(int device, size_t arena_bytes, gcv_ctx** out);
int gcv_ctx_destroy(gcv_ctx* ctx);
const char* gcv_last_error(const gcv_ctx* ctx);
int gcv_alloc(gcv_ctx* ctx, size_t bytes, void** dptr);
int gcv_free(gcv_ctx* ctx, void* dptr);
int gcv_arena_stats(const gcv_ctx* ctx, size_t* used, size_t* capacity);
int gcv_copy_h2d(gcv_ctx* ctx, void* dst, const void* src, size_t bytes);
int gcv_copy_d2h(gcv_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Device-to-device copy by the library's streaming kernel (16-byte aligned
 * pointers and size); enqueued, no sync. Used to measure the HBM ceiling. */
int gcv_copy_d2d(gcv_ctx* ctx, void* dst, const void* src, size_t bytes);
/* VALU rate calibration (no counterpart in the reference; a measurement
 * utility beside gcv_copy_d2d): n_blocks x 256 lanes each run 8 independent
 * chains of `iters` operations of one kind (0: FP32 FMA, 1: FP64 FMA,
 * 2: FP32 square root) and write one double to dbuf[n_blocks * 256].
 * Enqueued, no sync; timed by the profiling API under the names
 * k_valu_fma32 / k_valu_fma64 / k_valu_sqrt32. */
int gcv_valu_calibration(gcv_ctx* ctx, int kind, double* dbuf, int64_t n_blocks, int iters);
int gcv_sync(gcv_ctx* ctx);
int gcv_device_info(gcv_ctx* ctx, int* n_cu, char* name, int name_len);
/* The context's HIP stream (a hipStream_t), for kernels a caller compiles
 * itself, such as the header-only template surface (include/gcurval/gcurval.h)
 * instantiated on user curve plug-ins. No counterpart in the reference, whose
 * kernels run on the default stream. */
int gcv_ctx_stream(gcv_ctx* ctx, void** stream);

/* ---- timing: HIP events on the context stream ---- */
int gcv_timer_start(gcv_ctx* ctx);
int gcv_timer_stop(gcv_ctx* ctx, double* ms); /* synchronises */
/* Per-kernel launch timing (events around every launch while enabled). */
int gcv_profile_enable(gcv_ctx* ctx, int on);
int gcv_profile_reset(gcv_ctx* ctx);
/* Sum of durations and launch count of kernels whose name starts with prefix. */
int gcv_profile_query(gcv_ctx* ctx, const char* prefix, double* total_ms, int64_t* count);

/* ---- curve batches (device pointers, stay owned by the caller and must
 * outlive the table) ---- */

/* Ellipses x = a cos t, y = b sin t, t in [0, 2 pi], one knot interval each
 * (samples/sample1_ellipse/ellipse.h:17-65). a, b: device double[n_curve]. */
int gcv_table_create_ellipse(gcv_ctx* ctx, int n_curve, const double* d_a, const double* d_b,
                             const gcv_opts* opts, gcv_table** out);

/* Clamped B-splines of degree 1..5 in 2-D/3-D, mixed degrees allowed
 * (samples/sample2_bspline/structs.h:6-25, 126-163). Curve c has knots
 * knots[knot_off[c] .. knot_off[c+1]) and control points
 * ctrl[(ctrl_off[c] + j) * ctrl_stride + d]; its knot intervals are the spans
 * [t_{s}, t_{s+1}], s = p .. n_knots-p-2, i.e. n_knots - 2p - 1 intervals.
 * All arrays are device pointers. */
typedef struct gcv_bspline_batch {
    int n_curve;
    int ctrl_stride;          /* 2 or 3 doubles per control point */
    const int32_t* degree;    /* [n_curve] */
    const int32_t* dim;       /* [n_curve], 2 or 3 (<= ctrl_stride) */
    const int64_t* knot_off;  /* [n_curve+1] */
    const int64_t* ctrl_off;  /* [n_curve+1] */
    const double* knots;
    const double* ctrl;
} gcv_bspline_batch;

int gcv_table_create_bspline(gcv_ctx* ctx, const gcv_bspline_batch* batch, const gcv_opts* opts,
                             gcv_table** out);

/* Enqueue the table build: ArclengthCalculateCuda::createArclengthTable
 * (gcurval/detail/cuda/arclength_calculate.h:185-210). No host sync, no
 * allocation; may be called repeatedly. */
int gcv_table_build(gcv_table* t);

/* Uniform build with 2^log2_roots roots per knot interval (0..6); the plain
 * build is log2_roots = log2 n_subdivided. Every root is tested with the
 * plain level-0 tolerance (tolerance / n_subdivided), halving per level;
 * in roots wider than the plain ones, a segment no wider than a plain root
 * is also accepted once the halves-versus-whole difference is within 4
 * units of rounding of its length (the test cannot resolve less).
 * d_cost (optional, device int32 [n_intervals]) receives the quadratures
 * each interval used. */
int gcv_table_build_uniform(gcv_table* t, int log2_roots, int32_t* d_cost);

/* Seeded build: knot interval i starts from 2^k_i roots, d_k device int8
 * [n_intervals] with k_i in 0..6 (for example predicted by the MLP); every
 * root still passes the plain adaptive error test (tolerance / n_subdivided
 * at level 0, halving per level, with the rounding floor above), so
 * refinement falls back to the plain method where the seed is too coarse.
 * Segments wider than a plain root are also refined where their node speeds
 * dip, and segments wider than 1/8 of their interval must agree with a
 * higher-order rule (GL-5 for GL-3 builds, GL-8 for GL-5 builds). No
 * counterpart in the reference (its pre-split is the fixed N_SUBDIVIDED of
 * gcurval/config.h:20-24). d_cost as above. */
int gcv_table_build_seeded(gcv_table* t, const int8_t* d_k, int32_t* d_cost);

/* Safety floor of the last seeded build (device int8 [n_intervals]): the
 * smallest k an interval is allowed, from the ratio of its minimum to maximum
 * speed (near-zero speed inside a wide root can fool the error test). A
 * seeded build uses max(k_i, floor_i). */
int gcv_table_seed_floor(gcv_table* t, const int8_t** d_floor);

/* Root layout of the last build: seeded flag, roots per interval (log2) of
 * a uniform build (interval i's roots at i << log2_roots in the root-prefix
 * view), and for a seeded build the device log2 root count k_i of every
 * interval (int8 [n_intervals]) and the device first root of every interval
 * (int64 [n_intervals]; the roots are dense in interval order, interval i's
 * 2^k_i roots at root_off[i]). Any output may be null. */
int gcv_table_root_layout(gcv_table* t, int* seeded, int* log2_roots, const int8_t** d_root_k,
                          const int64_t** d_root_off);

int gcv_table_destroy(gcv_table* t);
int gcv_table_info(const gcv_table* t, int* n_curve, int64_t* n_intervals, int* n_subdivided);

/* Per-curve lengths to host: getAllCurveArclengths / getArclength
 * (gcurval/detail/cuda/arclength_calculate.h:212-222, 372-406). Synchronises. */
int gcv_table_lengths(gcv_table* t, double* host_out);

/* Device views of the table (valid until the next build or destroy):
 * per-interval FP64 cumulative lengths within each curve, per-root inclusive
 * FP64 prefix within each interval (normalised units), per-interval refined
 * root masks, per-curve lengths, interval offsets [n_curve+1]. */
int gcv_table_device_views(gcv_table* t, const double** int_cum, const double** root_prefix,
                           const uint64_t** deep_mask, const double** curve_len, const int64_t** int_off);

/* Build statistics of the last build: quadratures evaluated, refined roots,
 * maximum task-stack depth, stack overflow flag (must be 0), refinement tasks,
 * leaves that still failed the error test when the depth cap stopped them,
 * tiles whose refinement outgrew the small task stack of the first pass
 * (256 entries; their intervals are re-run with the full stack, and the
 * counts include both passes' work), 64-root wave tiles, and refinement
 * passes (wave passes over popped tasks). Any pointer may be NULL. */
int gcv_table_stats(gcv_table* t, uint64_t* quadratures, uint64_t* deep_roots, uint64_t* max_stack,
                    uint64_t* overflow, uint64_t* tasks, uint64_t* depth_capped, uint64_t* reruns,
                    uint64_t* tiles, uint64_t* passes);

/* Arc-length sampling: arclengthReverse / kn_initSearchTask +
 * kn_arclengthReverseDfsFast_Paper (gcurval/detail/cuda/arclength_calculate.h:224-231,
 * 408-505; kernel/reverse.inl.cuh:32-99, 216-341).
 *   d_sample_off: device int64[n_curve+1], samples of curve c are
 *                 [off[c], off[c+1]) (the partial sum of n_ptrs_for_curves).
 *                 The offsets must be non-decreasing with off[0] = 0 and
 *                 off[n_curve] = n_total (they stay on the device, so the
 *                 call does not check them): other offsets give unspecified
 *                 parameters, but no out-of-bounds access and no hang.
 *   d_pos:        device double[n_total] positions, or NULL for uniform
 *                 positions j/(N_c-1) (uniform_closed=1) or j/N_c (0).
 *   is_relative:  positions are fractions of the curve length (else arc lengths).
 *   d_out_u:      device double[n_total] curve parameters u (t for ellipses).
 *   d_out_seg:    optional device int32[n_total] local knot-interval index.
 * Enqueued on the context stream; no host sync. */
int gcv_sample(gcv_table* t, const int64_t* d_sample_off, int64_t n_total, const double* d_pos,
               int uniform_closed, int is_relative, double* d_out_u, int32_t* d_out_seg);

/* gcv_sample that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as gcv_sample (same kernels), but the per-256-sample block maps
 * live in a static device array of the library (8 MiB per device, held until
 * the process ends) instead of the table: at most 268,435,456 samples
 * (2^20 blocks of 256), larger n_total returns GCV_ERR_INVALID. Every call on
 * one GPU shares that array, so calls that overlap on one GPU (other
 * streams, other contexts) race; calls on one context's stream are ordered.
 * No profiling events are recorded for it. */
int gcv_sample_noalloc(gcv_table* t, const int64_t* d_sample_off, int64_t n_total, const double* d_pos,
                       int uniform_closed, int is_relative, double* d_out_u, int32_t* d_out_seg);

/* Curve points at sampled parameters: the gcurvalReverseReturnPoints variant
 * of arclengthReverse (kn_toPoints, gcurval/detail/cuda/kernel/reverse.inl.cuh:344-363;
 * cuda/arclength_calculate.h:458-500). d_u and d_seg are gcv_sample's outputs
 * for the same d_sample_off (d_seg may be NULL for ellipses); d_out: device
 * double [n_total][out_dim], raw coordinates, out_dim 2..3 (at least the
 * batch's ctrl_stride for B-splines; missing components are 0). */
int gcv_points(gcv_table* t, const int64_t* d_sample_off, int64_t n_total, const double* d_u,
               const int32_t* d_seg, int out_dim, double* d_out);

/* ---- learned seed: a 16-64-64-4 MLP on BF16 matrix cores (FP32 accumulate)
 * predicting, per knot interval, log2 of the initial root count (output 0),
 * whether every root of that seed passes the level-0 error test (output 1),
 * the scaled log2 of the worst quintic Newton-start error over its roots
 * (output 2) and whether that start may go straight to the FP64 step
 * (output 3). No counterpart in the reference; the adaptive error test still
 * verifies every root (see gcv_table_build_seeded) and the sampler every
 * FP64 step (see gcv_table_set_newton_start). ---- */
typedef struct gcv_mlp gcv_mlp;
int gcv_mlp_create(gcv_ctx* ctx, gcv_mlp** out);
int gcv_mlp_destroy(gcv_mlp* m);
/* Host FP32 parameters in the order W1[16][64], b1[64], W2[64][64], b2[64],
 * W3[64][4], b3[4] (5508 values) and the feature standardisation
 * (x - mean) * istd (16 + 16). Setting resets the optimiser state. */
int gcv_mlp_set_params(gcv_mlp* m, const float* params, const float* mean, const float* istd);
int gcv_mlp_get_params(gcv_mlp* m, float* params, float* mean, float* istd);
/* Raw 16 features of every knot interval of a table (device float
 * [n_intervals][16]); runs the table's setup kernels first (the build
 * statistics of gcv_table_stats are left as they are). */
int gcv_mlp_features(gcv_table* t, float* d_out);
/* Predicted k (device int8 [n_intervals]) and optionally the raw outputs
 * (device float [n_intervals][4]); hedge=1 adds one level where the pass head
 * predicts a level-0 failure. */
int gcv_mlp_predict_k(gcv_mlp* m, gcv_table* t, int hedge, int8_t* d_k, float* d_raw);
/* Setup + prediction + seeded build in one enqueue (no host sync); also keeps
 * the Newton-start decisions (output 3) for the table's sampling. */
int gcv_table_build_mlp(gcv_table* t, gcv_mlp* m, int hedge, int32_t* d_cost);
/* k predicted by the last gcv_table_build_mlp (device int8 [n_intervals]). */
int gcv_table_seed_k(gcv_table* t, const int8_t** d_k);
/* Forward pass over n standardised feature rows given as BF16 bits (device
 * uint16 [n][16]) into device float [n][4]. */
int gcv_mlp_forward(gcv_mlp* m, const uint16_t* d_x, int64_t n, float* d_out);
/* Loss and parameter gradients (host float [5508]) of
 * mean((out0 - y0)^2 + w_pass BCE(sigmoid(out1), y1) + (out2 - y2)^2
 *      + w_pass BCE(sigmoid(out3), y3)) over n rows, targets device float
 * [n][4]; a NaN target leaves its term out. Synchronises. */
int gcv_mlp_grad(gcv_mlp* m, const uint16_t* d_x, const float* d_y, int64_t n, float w_pass, float* host_grad,
                 float* host_loss);
/* One Adam step (beta 0.9 / 0.999) on that loss; host_loss optional (syncs). */
int gcv_mlp_train_step(gcv_mlp* m, const uint16_t* d_x, const float* d_y, int64_t n, float lr, float w_pass,
                       float* host_loss);

/* ---- graphs: gcv_graph_begin starts capturing the context stream; the
 * library calls made until gcv_graph_end (builds, learned builds, sampling)
 * are recorded, not run, and gcv_graph_launch replays them with one launch.
 * Calls that synchronise (copies to the host, lengths, statistics, gcv_sync)
 * fail with GCV_ERR_STATE while capturing; run a call once before capturing
 * it so that its lazily sized buffers exist. The graph keeps the tables and
 * buffers it was recorded with: they must outlive it. Replaces the
 * reference's per-level host loop with its synchronisations
 * (gcurval/detail/cuda/arclength_calculate.h:259-323). ---- */
typedef struct gcv_graph gcv_graph;
int gcv_graph_begin(gcv_ctx* ctx);
int gcv_graph_end(gcv_ctx* ctx, gcv_graph** out);
int gcv_graph_launch(gcv_graph* g);
int gcv_graph_destroy(gcv_graph* g);

/* ---- autotuner: builds the table with each launch shape of the persistent
 * subdivision kernel (1, 2 or 4 waves per block x 100, 50 or 25 % of the
 * resident-block limit), `reps` timed builds each, keeps the fastest for
 * this table and rebuilds with it. m: time learned-seed builds with this
 * predictor (NULL: plain builds). No counterpart in the reference (fixed
 * launch shapes). ---- */
typedef struct gcv_tune {
    int sub_waves_per_block;
    int sub_grid_pct;
    double best_build_ms;
    double default_build_ms; /* 4 waves per block, full grid, one stream */
    int n_tried;
    int bin_streams;         /* learned builds: streams the per-class passes are spread over */
} gcv_tune;
int gcv_table_autotune(gcv_table* t, gcv_mlp* m, int reps, gcv_tune* out);
/* Set the launch shape directly (0 waves per block: the default). bin_streams:
   1 = a build's per-class passes (coefficients, subdivision launches) on the
   context stream, 2 or 3 = spread over that many streams and joined before
   the build returns, 0 = the default (2 for B-spline batches of several
   degree/dimension classes, else 1). */
int gcv_table_set_launch(gcv_table* t, int sub_waves_per_block, int sub_grid_pct, int bin_streams);
/* Cumulative-scan kernel of B-spline builds (no counterpart in the reference,
   which scans per curve, kernel/generic.inl.cuh:117-137): 0 = by curve length
   (a wave per curve when every curve has at most 8,192 intervals and they
   average at least 64, else the single-pass look-back scan over all
   intervals), 1 = a wave per curve,
   2 = the look-back scan. Same results either way up to FP64 rounding; each
   is bit-reproducible from build to build. */
int gcv_table_set_scan(gcv_table* t, int mode);

/* Newton start of the B-spline sampler in FP32 GL-3 builds that take the
   final FP64 step (no counterpart in the reference, whose sampler bisects
   down its tree, kernel/reverse.inl.cuh:216-341; other tables always use
   GCV_START_PLAIN):
     GCV_START_PLAIN    (default) cubic Hermite start of the inverse map,
                        FP32 Newton steps, then the FP64 step;
     GCV_START_LEARNED  on tables built by gcv_table_build_mlp, the
                        intervals whose predictor output 3 is positive start
                        from the quintic Hermite interpolant (end speeds and
                        speed slopes) and take the FP64 step directly when
                        every sample of a wavefront is on such an interval
                        and on an unrefined root; a sample whose FP64 step
                        moves more than 2^-12 of its root falls back to the
                        plain start and steps. Tables built otherwise: as
                        GCV_START_PLAIN;
     GCV_START_QUINTIC  the quintic start with that check on every unrefined
                        root. */
#define GCV_START_PLAIN 0
#define GCV_START_LEARNED 1
#define GCV_START_QUINTIC 2
int gcv_table_set_newton_start(gcv_table* t, int mode);
/* Training labels of the Newton-start outputs on the table as built (FP32
   GL-3 B-spline tables; no counterpart in the reference): d_out[i] = log2 of
   the worst quintic-start error over interval i's roots, relative to root
   width, at 1/4, 1/2 and 3/4 of each root's length, against the converged
   FP64 Newton solution; a refined root counts as error 1. float [n_intervals]. */
int gcv_table_start_labels(gcv_table* t, float* d_out);
/* Diagnostic of the following gcv_sample calls (no counterpart in the
   reference): per sample, FP32 Newton steps | FP64 steps << 4 (int8
   [n_total]); null turns it off. */
int gcv_table_set_iters_out(gcv_table* t, int8_t* d_iters);

/* Host-side helpers that run without a GPU (used by the CPU test suite). */
/* Power-basis derivative coefficients of one B-spline span, as computed by
 * the setup kernel (out[d*p + j] multiplies x^j, x in [0,1] across the span). */
int gcv_host_bspline_span_poly(int p, int D, int s, const double* knots, const double* ctrl, int ctrl_stride,
                               double scale_inv, double* out, double* u0, double* h);
/* Library version string. */
const char* gcv_version(void);

#ifdef __cplusplus
}
#endif

#endif /* GCURVAL_C_H */
