/*
 * librcsrhip.so — the MI355X (gfx950) device arm of rcsr, as a C ABI.
 *
 * Drop-in for the reference engine's librcsrcuda.so on the batched
 * evaluation and batched LM/IRLS constant-fitting path.  Every entry point
 * below that carries a reference citation has that entry point's name,
 * signature, argument meaning and error behaviour (reference paths are
 * relative to the reference repository's my-version/src; the Rust binding of
 * each is in gpu.rs:668-825).  Entry points marked NEW extend the ABI for
 * the heterogeneous batch, the implicit relation, the plain-FP64 switch and
 * the test hooks; the reference has no counterpart for them.
 *
 * Conventions (eval.cu:3135-4893):
 *  - a context owns all device memory and is freed by rc_destroy;
 *  - host arrays are owned by the caller and must stay valid for the call;
 *  - calls are synchronous from the caller's view;
 *  - rc_create returns NULL on failure, setters return 0 / -1, batch calls
 *    return a kernel time in milliseconds that is NEGATIVE on failure (the
 *    caller then discards every output of that call);
 *  - per-candidate fit statuses: 0 ok, 1 ill-conditioned / no finite step,
 *    2 non-finite rows, 4 rank-deficient, 5 recovered, 6 no improvement,
 *    7 start unusable, 8 recovered with skipped gradient rows
 *    (fit/gpu.rs:97-105; eval.cu:2583-2605);
 *  - a context is driven by one host thread.
 *
 * Data layout: columns are column-major, cols[var * n + row].
 */
#ifndef RCSR_HIP_H
#define RCSR_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Host-side ledger of one batch in milliseconds (eval.cu:95-107 = gpu.rs:570). */
typedef struct RcStats {
    double kernel_ms;
    double launch_ms;
    double sync_ms;
    double alloc_ms;
    double h2d_ms;
    double d2h_ms;
    double wall_ms;
    unsigned long long h2d_bytes;
    unsigned long long d2h_bytes;
    int batch;
    int failed;
} RcStats;

/* Arena / fit-region accounting (eval.cu:119-136 = gpu.rs:591). */
typedef struct RcArenaReport {
    int async_copy;
    char reason[96];
    unsigned long long pinned_bytes;
    unsigned long long alloc_calls;
    unsigned long long steady_alloc_calls;
    unsigned long long overlapped_uploads;
    unsigned long long batches;
    unsigned long long fs_alloc_calls;
    unsigned long long fs_growths;
    unsigned long long fs_steady_alloc_calls;
    unsigned long long fs_batches;
    unsigned long long fs_dev_bytes;
    unsigned long long fs_pinned_bytes;
} RcArenaReport;

/* eval.cu:4529-4535 RcMaxOut: one segment of rc_max_scan_batch. */
typedef struct RcMaxOut {
    double vmax; /* exact max |r|, NaN when status != 0 */
    double lo;   /* max of the per-row lower bounds (a level was requested) */
    double hi;   /* min of the per-row upper bounds */
    int row;     /* lowest-index maximiser, -1 when status != 0 */
    int status;  /* 0 ok; 2 a row was not finite */
} RcMaxOut;

/* ---- context and data -------------------------------------------------- */

/* eval.cu:3137 rc_create. Uploads n x v f32 columns and the f32 target. */
void *rc_create(const float *cols, int n, int v, const float *y);

/* eval.cu:3221 rc_destroy. */
void rc_destroy(void *ctx);

/* eval.cu:3882 rc_set_fit_data: exact f64 columns, target and optional base
   row weights (NULL = unweighted) for the fit's exact passes. */
int rc_set_fit_data(void *ctx, const double *cols, const double *y, const double *w);

/* eval.cu:3935 rc_set_score_weights: base weights shared with the fit;
   NULL clears them. */
int rc_set_score_weights(void *ctx, const double *w);

/* eval.cu:3955 rc_set_fit_sign: per-row label sign t = 2l - 1 of the
   resident rows for the logistic fit arm (metric 5) and the classification
   scores (10-14); NULL clears it, after which those launches return -1. */
int rc_set_fit_sign(void *ctx, const double *sign);

/* ---- evaluation -------------------------------------------------------- */

/* eval.cu:3268 rc_set_programs: upload a packed postfix batch (words
   HALT-terminated; off has nprog+1 word offsets). */
int rc_set_programs(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog);

/* eval.cu:3344 rc_eval_submit: enqueue the batch's loss sums (loss: 0
   squared, 1 log-squashed, 4 absolute).  A non-finite row or |pred| > 1e30
   adds 1e30 (eval.cu:759-793). */
int rc_eval_submit(void *ctx, int loss);

/* eval.cu:3395 rc_eval_collect: wait, copy nprog f32 sums, return ms. */
float rc_eval_collect(void *ctx, float *out);

/* eval.cu:3429 rc_eval: submit + collect. */
float rc_eval(void *ctx, float *out, int loss);

/* No reference counterpart: for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc).  The same losses as rc_set_programs(prog, off, nprog) followed
   by rc_eval(out, loss), bit for bit, but the call allocates nothing: the
   translated streams, partial sums and outputs live in the library's static
   device array of 32 MiB per GPU (held until the process ends), and only
   the derived columns the context already holds are used.  The context's
   current batch is left as it was.  Calls that overlap on one GPU race on
   that array.  Returns the device time in ms; -2 when the batch does not fit
   the array (about nprog * ceil(n / 4096) * 12 bytes of partial sums plus
   the streams), -1 on any other failure or while a submitted batch is
   pending. */
float rc_eval_static(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog,
                     float *out, int loss);

/* eval.cu:3438 rc_predict: raw P x N f32 predictions, row-major. */
int rc_predict(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog,
               float *out);

/* eval.cu:3803 rc_score_batch: native score objectives with f64
   accumulation (kind: 0 squared, 1 absolute, 2 log-squashed, 3 signed,
   4 maximum, 7/8/9 AIC/AICc/BIC); non-finite -> +inf and status 2. */
float rc_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off, int nprog,
                     int kind, const int *pcount, double *out, int *status);

/* eval.cu:3514 rc_pool_reserve / 3550 rc_pool_upload / 3565 rc_pool_eval /
   3823 rc_pool_score / 3864 rc_pool_hash: the fixed-width genome pool
   (slot p starts at p * width and ends at its HALT word). */
int rc_pool_reserve(void *ctx, int nprog, int width);
int rc_pool_upload(void *ctx, int which, const unsigned int *prog, int nprog);
float rc_pool_eval(void *ctx, int which, int nprog, int width, float *out, int loss);
float rc_pool_score(void *ctx, int which, int nprog, int width, int kind, const int *pcount,
                    double *out, int *status);
int rc_pool_hash(void *ctx, int which, int nprog, int width, unsigned long long *out);

/* One generation of the flat postfix variation (postfix.rs:500-893, host
   code in the reference; Philox4x32-10 of genome.rs:214-300) computed on the
   device from pool arena `from` into the other arena: per island the best
   parent survives in slot 0, islands with restart[i] != 0 are refilled with
   random programs, every other slot is a child keyed by (seed, island,
   generation, slot, attempt).  ops holds nops (code, arity) pairs; err holds
   the parents' errors (NaN ranks last).  The optional outputs receive each
   slot's length in words, node count, FNV-1a hash and inherited error (NaN
   for new programs).  Returns 0, or -1 on an invalid argument or a device
   error; the pool width must be at most 128 words. */
int rc_pool_vary(void *ctx, int from, int islands, int pop, unsigned int seed,
                 unsigned int generation, int max_complexity, int nvars, int init_depth,
                 int dedup, const unsigned int *ops, int nops, const double *err,
                 const unsigned char *restart, unsigned short *out_len, unsigned short *out_cost,
                 unsigned long long *out_hash, double *out_err);
/* Copies the first nprog slots of arena `which` to out (nprog * width
   words).  Returns 0 or -1. */
int rc_pool_download(void *ctx, int which, unsigned int *out, int nprog);

/* ---- fit --------------------------------------------------------------- */

/* eval.cu:4001 rc_fit_batch: batched LM / IRLS constant fitting.  prog holds
   post-order 4-word records [op, kid0, kid1, slot] (fit.rs:1101-1196),
   off nprog+1 word offsets, pcount the constant count of each candidate,
   c0 nprog x kmax start constants; gauge (may be NULL) holds per candidate
   the (scale, offset) constant slots of its outer affine pair, -1 for none.
   metric: 0 squared, 1 log-squashed, 4 absolute (objective sum w*rho(r));
   5 logistic (eval.cu:683-705, 1569-1576: objective sum w ln(1 + exp(-t p))
   with the signs of rc_set_fit_sign, Fisher-weighted LM; -1 without signs);
   2 correlation 1 - |rho|, 3 signed correlation 1 - rho (eval.cu:2320-2448:
   the gauge pair is renormalised to (sign(a), 0) and frozen; objective the
   weighted Pearson form).  Outputs: constants (nprog x kmax), the fitted
   objective in f64, and the status.  kmax > 20 -> -1. */
float rc_fit_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                   const int *pcount, int nprog, int kmax, const double *c0, const int *gauge,
                   int iters, int metric, double *outc, double *outsse, int *outstatus);

/* eval.cu:4084 rc_fit_kernel_info: launch shape of the fit path for a batch
   whose largest constant count is kmax (threads per workgroup, LDS bytes per
   workgroup, constant cap); -1 when kmax > 20. */
int rc_fit_kernel_info(int kmax, int *threads, int *smem_bytes, int *max_constants);

/* ---- exact order, rank and maximum statistics --------------------------- */

/* Each call scores nsets x ntrials segments: segment set * ntrials + trial
   evaluates set `set`'s post-order fit records (the rc_fit_batch program
   form, pcount constants) over the resident f64 rows with the constants
   c0[(set * ntrials + trial) * kmax ...].  kmax <= 20, at most 96 nodes,
   nsets * ntrials * n < 2^31.  They return the device time in ms, negative
   on failure (the caller then uses its exact CPU arm). */

/* eval.cu:4743 rc_order_stat_score_batch: the median (what 0) or the
   interquartile mean (what 1, ranks [n/4, 3n/4), all rows below four) of
   |pred - y| per segment; status 2 when a row is not finite (value NaN). */
float rc_order_stat_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                                const int *pcount, int nsets, int ntrials, int kmax,
                                const double *c0, int what, double *out_value, int *out_status);

/* eval.cu:4270 rc_rank_score_batch: exact rank sums against the target.
   With A and B the doubled centred average ranks (2 r - (n + 1), ties share
   their average) of the prediction and of y: out[4 seg ..] = Sigma A B and
   Sigma A A as two's-complement 128-bit (hi, lo) pairs, out_bb = Sigma B B;
   status 2 when a prediction is not finite; n <= 1 or a non-finite target
   fails the call. */
float rc_rank_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                          const int *pcount, int nsets, int ntrials, int kmax, const double *c0,
                          unsigned long long *out, int *out_status, unsigned long long *out_bb);

/* eval.cu:4852 rc_rank_debug_b: the target's cached B per row. */
int rc_rank_debug_b(void *ctx, long long *out);

/* eval.cu:4651 rc_max_scan_batch: per segment, the exact max |pred - y| and
   its lowest row; with coords[seg] a constant slot (< pcount) and levels[seg]
   a finite level t >= 0, also the interval [lo, hi] of steps delta along
   that constant with |r_i + (d r_i / d c) delta| <= t on every row (a row
   whose derivative is 0 and whose |r_i| > t makes it empty: lo = +inf,
   hi = -inf).  coords / levels may be NULL (maximum only). */
float rc_max_scan_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                        const int *pcount, int nsets, int ntrials, int kmax, const double *c0,
                        const int *coords, const double *levels, RcMaxOut *out);

/* NEW: select the device arithmetic of rc_fit_batch for this context.
   0 = mixed precision (FP32 interpreter and adjoint, FP32-MFMA split Gram,
   FP64 solve / acceptance / IRLS weights; the default); 1 = plain FP64 device
   path (every row pass in f64).  Returns the previous mode or -1. */
int rc_set_fit_mode(void *ctx, int mode);

/* NEW, no counterpart in the reference: the launch shapes of this context,
   chosen by the autotuner (rcsr_hip.autotune) or the caller; value 0 restores
   the built-in choice.  A candidate's results depend on its own program and
   rows and on these shapes, never on the rest of its batch.
     RC_TUNE_FIT_TILE_ROWS      rows per fit work item: a multiple of 256 in
                                [256, 4194304] (built-in 8192)
     RC_TUNE_FIT_ROWS_PER_LANE  rows per lane of the fit VM assembly: 4 where
                                the program fits the 4-row files (built-in) or 2
     RC_TUNE_EVAL_TILE_ROWS     rows per evaluation tile: a multiple of 256 in
                                [256, 4194304] (built-in 4096)
   rc_set_tuning returns the previous value, or -1 for an unknown key or a
   value out of range; rc_get_tuning returns the current value or -1. */
#define RC_TUNE_FIT_TILE_ROWS 1
#define RC_TUNE_FIT_ROWS_PER_LANE 2
#define RC_TUNE_EVAL_TILE_ROWS 3
int rc_set_tuning(void *ctx, int key, int value);
int rc_get_tuning(void *ctx, int key);

/* NEW: add a resident row set for heterogeneous batches; returns its id
   (the rc_create set is id 0) or -1.  cols/y are f32 column-major, cols64/y64
   their exact f64 twins, w optional base weights. */
int rc_add_rows(void *ctx, const float *cols, int n, int v, const float *y,
                const double *cols64, const double *y64, const double *w);

/* NEW: heterogeneous batch in one launch sequence: per-candidate row set,
   metric (0 / 1 / 4, or 5 on row set 0) and the shared iteration count.  Same outputs as
   rc_fit_batch. */
float rc_fit_batch_hetero(void *ctx, const unsigned int *prog, const unsigned int *off,
                          const int *pcount, int nprog, int kmax, const double *c0,
                          const int *rowset, const int *metric, int iters, double *outc,
                          double *outsse, int *outstatus);

/* NEW (test hook): per-row f32 prediction, constant gradient (kmax wide) and
   input gradient (v wide) of candidate `cand` of a fit batch at constants c,
   for the first `rows` rows of row set 0.  Returns 0 or -1. */
int rc_fit_jacobian(void *ctx, const unsigned int *prog, const unsigned int *off,
                    const int *pcount, int nprog, int kmax, const double *c, int cand, int rows,
                    float *out_pred, float *out_jac, float *out_gx);

/* NEW: implicit relation score, mean over rows of |f| / ||grad_x f|| (metric
   0) or |f| (metric 1) for fit-format programs at constants c, computed with
   the adjoint input-gradient sweep; +inf and status 2 when a row is
   undefined (implicit.rs:59-155). */
float rc_implicit_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                              const int *pcount, int nprog, int kmax, const double *c,
                              int metric, double *out, int *status);

/* NEW: implicit relation fit: LM on the signed geometric residual
   f / ||grad_x f|| with the central-difference Jacobian of implicit.rs:277-409;
   outsse is the final sum of squared geometric residuals. */
float rc_implicit_fit_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                            const int *pcount, int nprog, int kmax, const double *c0, int iters,
                            double *outc, double *outsse, int *outstatus);

/* ---- telemetry --------------------------------------------------------- */

/* NEW: per-pass device time of the fit path, for roofline accounting. */
typedef struct RcFitReport {
    double asm_ms;      /* forward + adjoint + MFMA Gram passes */
    double trial_ms;    /* f32 trial-loss passes */
    double exact_ms;    /* f64 exact-loss passes (re-baseline, polish) */
    double ctrl_ms;     /* reductions, solves, acceptance */
    int asm_launches;
    int trial_launches;
    int exact_launches;
    int ctrl_launches;
} RcFitReport;

/* NEW: enable (1) or disable (0) per-pass timing of later fits on this
   context; while enabled the launch sequence runs eagerly with event pairs
   around every pass instead of as one graph.  Resets the totals.  Returns the
   previous setting or -1. */
int rc_fit_profile(void *ctx, int enable);

/* NEW: totals since rc_fit_profile, and (when per_cand is non-NULL) five
   ints per candidate of the last fit: the assembly, trial and f64-exact
   passes it consumed, its LM accept/reject record (bit q set when the q-th
   LM decision accepted the trial, for q < 31) and the number of LM
   decisions.  Returns 0 or -1. */
int rc_fit_report(void *ctx, RcFitReport *out, int *per_cand, int nprog);

/* eval.cu:3836 rc_stats_begin / 3855 rc_last_stats. */
void rc_stats_begin(void *ctx);
int rc_last_stats(void *ctx, RcStats *out);

/* eval.cu:3469 rc_arena_report / 3494 rc_arena_force_sync. */
int rc_arena_report(void *ctx, RcArenaReport *out);
int rc_arena_force_sync(void *ctx, int sync);

/* ---- probe ------------------------------------------------------------- */

/* eval.cu:4886 rc_device_count / 4872 rc_device_name / 4863 rc_unified
   (probed by gpu.rs:921). */
int rc_device_count(void);
const char *rc_device_name(void);
int rc_unified(void);

#ifdef __cplusplus
}
#endif
#endif
