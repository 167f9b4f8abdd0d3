/*
 * CPU f64 oracle for the rcsr device arm — TEST INFRASTRUCTURE ONLY.
 *
 * This library is a restatement, in C++, of the reference engine's exact
 * CPU paths that check the device arm: the tree evaluator with its domain
 * guards, the flat post-order form, the adjoint sweep, the Levenberg-Marquardt
 * constant fitter, the L1/log-squashed IRLS fitter, the weighted metric
 * scorers and the implicit-relation distance/fitter.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as the checker; the product library never links or calls it.
 *
 * Tree format across this ABI: a pre-order node list.  `ops[i]` is the
 * operator code (the device opcode numbering for every operator the device
 * knows, plus OR_ASINH/OR_ACOSH/OR_ATANH), `vals[i]` is the constant value of
 * an OR_CONST node or the column index of an OR_VAR node and is ignored
 * otherwise.  Arity follows from the operator.
 *
 * Row data is row-major: x[row * nvars + var].  `w` may be NULL (unweighted).
 */
#ifndef RCSR_ORACLE_H
#define RCSR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_CONST = 0, OR_VAR = 1, OR_ADD = 2, OR_SUB = 3, OR_MUL = 4, OR_DIV = 5,
    OR_SIN = 6, OR_COS = 7, OR_TAN = 8, OR_EXP = 9, OR_LOG = 10, OR_SQRT = 11,
    OR_ABS = 12, OR_TANH = 13, OR_SIG = 14, OR_SQUARE = 15, OR_CUBE = 16,
    OR_NEG = 17, OR_GAUSS = 18, OR_STEP = 19, OR_POW = 20, OR_SINH = 21,
    OR_COSH = 22, OR_FLOOR = 23, OR_CEIL = 24, OR_ROUND = 25, OR_MOD = 26,
    OR_ATAN2 = 27, OR_ERF = 28, OR_ERFC = 29, OR_MIN = 30, OR_MAX = 31,
    OR_IF = 32, OR_SIGN = 33, OR_EQUAL = 34, OR_LESS = 35,
    OR_LESS_OR_EQUAL = 36, OR_GREATER = 37, OR_GREATER_OR_EQUAL = 38,
    OR_AND = 39, OR_OR = 40, OR_XOR = 41, OR_NOT = 42,
    OR_ASINH = 43, OR_ACOSH = 44, OR_ATANH = 45
};

/* Metric codes (the device fit codes): squared, log-squashed, absolute. */
enum { OR_METRIC_SQUARED = 0, OR_METRIC_LOG_SQUASHED = 1, OR_METRIC_ABSOLUTE = 4 };
/* The log-loss metric (the logistic arm's objective) of or_fit_metric /
   or_metric_score: mean ln(1 + exp(-t p)) with t from the target midpoint. */
enum { OR_METRIC_LOG_LOSS = 10 };

/* Arity of an operator code, -1 for an unknown code. */
int or_arity(int op);

/* expr::eval of one row. */
double or_eval(const int *ops, const double *vals, int nn, const double *x, int nx);

/* expr::erf / expr::erfc. */
double or_erf(double x);
double or_erfc(double x);

/* Number of constant slots of the flat form. */
int or_nconst(const int *ops, const double *vals, int nn);

/* Flat `sse` (fit.rs sse) of the tree's own constants. */
double or_sse(const int *ops, const double *vals, int nn, const double *x, int n,
              int nv, const double *y, const double *w);

/* Predictions of the flat (fit) form over n rows, the tree's own constants
   (fit.rs:1424-1440 eval_row per row). */
void or_predict_flat(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                     double *out);

/* Adjoint gradient of the root w.r.t. the constant slots (g has nconst
   entries) and w.r.t. the inputs (gx has nv entries); either may be NULL. */
void or_gradients(const int *ops, const double *vals, int nn, const double *x,
                  int nv, double *g, double *gx);

/* fit_weighted (exact LM, fit_lm): returns the weighted SSE, writes the
   fitted constants (slot order) to out_c. */
double or_fit_weighted(const int *ops, const double *vals, int nn,
                       const double *x, int n, int nv, const double *y,
                       const double *w, int iters, double *out_c);

/* fit_weighted_metric for squared / log-squashed / absolute. */
double or_fit_metric(const int *ops, const double *vals, int nn, const double *x,
                     int n, int nv, const double *y, const double *w, int iters,
                     int metric, double *out_c);

/* metrics::score_weighted of the tree's predictions (mean loss). */
double or_metric_score(const int *ops, const double *vals, int nn,
                       const double *x, int n, int nv, const double *y,
                       const double *w, int metric);

/* implicit::score (metric: 0 derivative distance, 1 plain residual). */
double or_implicit_score(const int *ops, const double *vals, int nn,
                         const double *x, int n, int nv, const double *w,
                         int metric);

/* implicit::signed_distance of one row; returns 0 and writes *d, or -1. */
int or_signed_distance(const int *ops, const double *vals, int nn,
                       const double *x, int nv, double *d);

/* implicit::fit: returns the final sum of squared geometric residuals and
   writes the constants (pre-order) to out_c. */
double or_implicit_fit(const int *ops, const double *vals, int nn,
                       const double *x, int n, int nv, const double *w,
                       int iters, double *out_c);

/* Batched helper for the CPU baseline and for large parity sweeps: fits
   `ntree` trees (concatenated pre-order lists; tree t occupies
   [tree_off[t], tree_off[t+1])) on `threads` host threads.  metric per tree
   (NULL = all squared).  out_c is ntree x kmax, out_loss ntree. */
void or_fit_batch(const int *ops, const double *vals, const int *tree_off,
                  int ntree, const int *metric, const double *x, int n, int nv,
                  const double *y, const double *w, int iters, int kmax,
                  int threads, double *out_c, double *out_loss);

/* Batched implicit fit, same layout; out_loss is the final geometric SSE. */
void or_implicit_fit_batch(const int *ops, const double *vals,
                           const int *tree_off, int ntree, const double *x,
                           int n, int nv, const double *w, int iters, int kmax,
                           int threads, double *out_c, double *out_loss);

/* The device correlation fit arm (metric 2: 1 - |rho|, 3: 1 - rho) with
   the affine gauge pair (slot indices, -1 for none), restated in f64 from
   eval.cu:2320-2448 with each step's system masked afresh: returns the
   objective, writes the constants (slot order) and the status label. */
double or_corr_fit(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                   const double *y, const double *w, int iters, int metric, int gauge_scale,
                   int gauge_offset, double *out_c, int *status);

/* The weighted correlation objective of the tree's own constants. */
double or_corr_score(const int *ops, const double *vals, int nn, const double *x, int n, int nv,
                     const double *y, const double *w, int metric);

/* Order, rank and maximum statistics of residual / prediction vectors. */

/* metrics::median_abs (metrics.rs:1193-1205). */
double or_median_abs(const double *r, int n);

/* metrics::interquartile_mean_abs (metrics.rs:1209-1223). */
double or_iqr_mean_abs(const double *r, int n);

/* metrics::centred_ranks_doubled (metrics.rs:1258-1286); -1 if not finite. */
int or_centred_ranks_doubled(const double *v, int n, long long *out);

/* RankSums::between (metrics.rs:1307-1323): out = ab_hi, ab_lo, aa_hi,
   aa_lo, bb_hi, bb_lo (two's-complement 128-bit). */
void or_rank_sums(const long long *a, const long long *b, int n, unsigned long long *out);

/* The exact maximum-error scan (eval.cu:4505-4600) of residuals r with
   derivatives b along one constant (b may be NULL: maximum only). */
void or_max_scan(const double *r, const double *b, int n, double level, double *vmax, int *row,
                 double *lo, double *hi, int *status);

#ifdef __cplusplus
}
#endif
#endif
