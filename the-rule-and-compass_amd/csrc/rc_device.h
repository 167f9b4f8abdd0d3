/*
 * Device-side definitions shared by the rcsr HIP kernels (gfx950).
 *
 * Opcode numbering is the reference's bytecode contract (eval.cu:34-83 =
 * gpu.rs:29-73) so programs are interchangeable across the C ABI.  The
 * per-operator arithmetic below follows the reference's device semantics for
 * f32 rows (rc_exec_row eval.cu:404-573, ft_forward 1173-1233) and the CPU
 * fitter's value_of for f64 rows (fit.rs:1211-1365).
 */
#ifndef RC_DEVICE_H
#define RC_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

enum {
    OP_CONST = 0, OP_VAR = 1, OP_ADD = 2, OP_SUB = 3, OP_MUL = 4, OP_DIV = 5,
    OP_SIN = 6, OP_COS = 7, OP_TAN = 8, OP_EXP = 9, OP_LOG = 10, OP_SQRT = 11,
    OP_ABS = 12, OP_TANH = 13, OP_SIG = 14, OP_SQUARE = 15, OP_CUBE = 16,
    OP_NEG = 17, OP_GAUSS = 18, OP_STEP = 19, OP_POW = 20, OP_SINH = 21,
    OP_COSH = 22, OP_FLOOR = 23, OP_CEIL = 24, OP_ROUND = 25, OP_MOD = 26,
    OP_ATAN2 = 27, OP_ERF = 28, OP_ERFC = 29, OP_MIN = 30, OP_MAX = 31,
    OP_IF = 32, OP_SIGN = 33, OP_EQUAL = 34, OP_LESS = 35,
    OP_LESS_OR_EQUAL = 36, OP_GREATER = 37, OP_GREATER_OR_EQUAL = 38,
    OP_AND = 39, OP_OR = 40, OP_XOR = 41, OP_NOT = 42,
    OP_HALT = 255
};

/* Internal instruction words (built on the host from the ABI's programs):
 *   bits  0..7   opcode
 *   bits  8..15  flags (tape / push bits, see rc_fit.hip)
 *   bits 16..31  argument (constant slot, column index)
 * A CONST instruction of the evaluator stream is followed by one word of f32
 * bits. */
#define RC_W_OP(w) ((w) & 0xffu)
#define RC_W_FLAGS(w) (((w) >> 8) & 0xffu)
#define RC_W_ARG(w) ((w) >> 16)

/* Fit loss codes (RC_LOSS_*, eval.cu:589-604). */
#define RC_LOSS_SQUARED 0
#define RC_LOSS_LOG_SQUASHED 1
#define RC_LOSS_CORRELATION 2
#define RC_LOSS_CORRELATION_SIGNED 3
#define RC_LOSS_ABSOLUTE 4
#define RC_LOSS_LOGISTIC 5

/* Native score codes (eval.cu:609-640). */
#define RC_SCORE_SQUARED 0
#define RC_SCORE_ABSOLUTE 1
#define RC_SCORE_LOG_SQUASHED 2
#define RC_SCORE_SIGNED 3
#define RC_SCORE_MAXIMUM 4
#define RC_SCORE_CORRELATION 5
#define RC_SCORE_CORRELATION_SIGNED 6
#define RC_SCORE_AIC 7
#define RC_SCORE_AICC 8
#define RC_SCORE_BIC 9
#define RC_SCORE_NONFINITE 2
/* Classification score codes (eval.cu:628-640): the prediction is a logit p,
   the label sign t = 2l - 1 comes from rc_set_fit_sign, rows are weighted by
   the score weights and the score is the weighted mean row loss. */
#define RC_SCORE_LOG_LOSS 10
#define RC_SCORE_BRIER 11
#define RC_SCORE_HINGE_SIGMOID 12
#define RC_SCORE_HINGE_ORDINARY 13
#define RC_SCORE_HINGE_SQUARED 14
#define RC_SCORE_ALL_ZERO_WEIGHT 4

#define RC_F32_NAN __builtin_nanf("")
#define RC_F64_NAN __builtin_nan("")

/* ---- f32 row semantics (the evaluator and the mixed-precision fitter). */

__device__ __forceinline__ float rc_logistic_f(float x) { return 1.0f / (1.0f + expf(-x)); }

/* Logistic without overflow for either sign of x (the two-branch form
   eval.cu:707-712 uses for classification scores). */
__device__ __forceinline__ float rc_logistic_stable_f(float x)
{
    if (x >= 0.0f) return 1.0f / (1.0f + expf(-x));
    const float e = expf(x);
    return e / (1.0f + e);
}

/* Classification row loss of logit p against label sign t (eval.cu:725-751):
   log loss log(1 + e^{-tp}) in its stable form, Brier (sigma(p) - l)^2 with
   l = (t + 1) / 2, hinge on the sigmoid margin, ordinary and squared hinge.
   Any other code is NaN. */
__device__ __forceinline__ float rc_class_row_loss_f32(float p, float t, int kind)
{
    const float u = t * p;
    switch (kind) {
        case RC_SCORE_LOG_LOSS: {
            const float a = fabsf(u);
            return ((u < 0.0f) ? -u : 0.0f) + log1pf(expf(-a));
        }
        case RC_SCORE_BRIER: {
            const float d = rc_logistic_stable_f(p) - 0.5f * (t + 1.0f);
            return d * d;
        }
        case RC_SCORE_HINGE_SIGMOID: {
            const float h = 1.0f - t * (2.0f * rc_logistic_stable_f(p) - 1.0f);
            return h > 0.0f ? h : 0.0f;
        }
        case RC_SCORE_HINGE_ORDINARY: {
            const float h = 1.0f - u;
            return h > 0.0f ? h : 0.0f;
        }
        case RC_SCORE_HINGE_SQUARED: {
            const float h = 1.0f - u;
            return h > 0.0f ? h * h : 0.0f;
        }
        default: return RC_F32_NAN;
    }
}

/* Unary operator on f32, the reference device semantics (eval.cu:519-563). */
__device__ __forceinline__ float rc_unary_f32(unsigned op, float a)
{
    switch (op) {
        case OP_SIN: return sinf(a);
        case OP_COS: return cosf(a);
        case OP_TAN: { float t = tanf(a); return isfinite(t) ? t : RC_F32_NAN; }
        case OP_EXP: return expf(a);
        case OP_LOG: return (a > 0.0f) ? logf(a) : RC_F32_NAN;
        case OP_SQRT: return (a >= 0.0f) ? sqrtf(a) : RC_F32_NAN;
        case OP_ABS: return fabsf(a);
        case OP_TANH: return tanhf(a);
        case OP_SIG: return rc_logistic_f(a);
        case OP_SQUARE: return a * a;
        case OP_CUBE: return a * a * a;
        case OP_NEG: return -a;
        case OP_GAUSS: return expf(-a * a);
        case OP_STEP: return a > 0.0f ? 1.0f : 0.0f;
        case OP_SINH: return sinhf(a);
        case OP_COSH: return coshf(a);
        case OP_FLOOR: return floorf(a);
        case OP_CEIL: return ceilf(a);
        case OP_ROUND: return roundf(a);
        case OP_ERF: return (a == 0.0f) ? 0.0f : erff(a);
        case OP_ERFC: return erfcf(a);
        case OP_SIGN:
            return isnan(a) ? RC_F32_NAN : (a > 0.0f ? 1.0f : (a < 0.0f ? -1.0f : 0.0f));
        case OP_NOT: return isnan(a) ? RC_F32_NAN : (a > 0.0f ? 0.0f : 1.0f);
        default: return RC_F32_NAN;
    }
}

/* Binary operator on f32 (eval.cu:428-512). `a` is the left operand. */
__device__ __forceinline__ float rc_binary_f32(unsigned op, float a, float b)
{
    switch (op) {
        case OP_ADD: return a + b;
        case OP_SUB: return a - b;
        case OP_MUL: return a * b;
        case OP_DIV: return (fabsf(b) < 1e-12f) ? RC_F32_NAN : a / b;
        case OP_POW:
            return (a < 0.0f && fabsf(b - roundf(b)) > 1e-9f) ? RC_F32_NAN : powf(a, b);
        case OP_MOD: {
            /* Floored remainder with one range correction of the quotient's
               rounding step. */
            float r = RC_F32_NAN;
            if (isfinite(a) && isfinite(b) && fabsf(b) >= 1e-12f) {
                float q = a / b;
                if (isfinite(q)) {
                    r = fmaf(-b, floorf(q), a);
                    float ab = fabsf(b);
                    float rs = (b > 0.0f) ? r : -r;
                    if (rs < 0.0f) r += b;
                    else if (rs >= ab) r -= b;
                }
            }
            return r;
        }
        case OP_ATAN2: return (isnan(a) || isnan(b)) ? RC_F32_NAN : atan2f(a, b);
        case OP_MIN: return (isnan(a) || isnan(b)) ? RC_F32_NAN : fminf(a, b);
        case OP_MAX: return (isnan(a) || isnan(b)) ? RC_F32_NAN : fmaxf(a, b);
        case OP_EQUAL: return (isnan(a) || isnan(b)) ? RC_F32_NAN : (a == b ? 1.0f : 0.0f);
        case OP_LESS: return (isnan(a) || isnan(b)) ? RC_F32_NAN : (a < b ? 1.0f : 0.0f);
        case OP_LESS_OR_EQUAL: return (isnan(a) || isnan(b)) ? RC_F32_NAN : (a <= b ? 1.0f : 0.0f);
        case OP_GREATER: return (isnan(a) || isnan(b)) ? RC_F32_NAN : (a > b ? 1.0f : 0.0f);
        case OP_GREATER_OR_EQUAL:
            return (isnan(a) || isnan(b)) ? RC_F32_NAN : (a >= b ? 1.0f : 0.0f);
        case OP_AND:
            return (isnan(a) || isnan(b)) ? RC_F32_NAN : ((a > 0.0f && b > 0.0f) ? 1.0f : 0.0f);
        case OP_OR:
            return (isnan(a) || isnan(b)) ? RC_F32_NAN : ((a > 0.0f || b > 0.0f) ? 1.0f : 0.0f);
        case OP_XOR:
            return (isnan(a) || isnan(b)) ? RC_F32_NAN : (((a > 0.0f) != (b > 0.0f)) ? 1.0f : 0.0f);
        default: return RC_F32_NAN;
    }
}

__device__ __forceinline__ bool rc_is_binary(unsigned op)
{
    return (op >= OP_ADD && op <= OP_DIV) || op == OP_POW || op == OP_MOD || op == OP_ATAN2 ||
           op == OP_MIN || op == OP_MAX || (op >= OP_EQUAL && op <= OP_XOR);
}

/* ---- Per-row loss (rc_row_loss eval.cu:648-664) and IRLS weight
   (ft_row_weight 672-677). */
__device__ __forceinline__ double rc_row_rho(double r, int metric)
{
    if (metric == RC_LOSS_ABSOLUTE) return fabs(r);
    if (metric == RC_LOSS_LOG_SQUASHED) return log1p(r * r);
    return r * r;
}

__device__ __forceinline__ double rc_row_weight(double base, double r, int metric)
{
    if (metric == RC_LOSS_LOG_SQUASHED) return base * 2.0 / (1.0 + r * r);
    if (metric == RC_LOSS_ABSOLUTE) return base / fmax(fabs(r), 1e-6);
    return base;
}

/* ---- The logistic fit arm (eval.cu:683-705, 1569-1576).  The row's target
   slot holds the label sign t = +-1; the margin is u = t p. */
/* sigmoid(u) in the branch form that cannot overflow, clamped to
   [1e-9, 1 - 1e-9] so t / q and q (1 - q) stay finite. */
__device__ __forceinline__ double rc_logistic_q(double u)
{
    double q;
    if (u >= 0.0) {
        q = 1.0 / (1.0 + exp(-u));
    } else {
        const double e = exp(u);
        q = e / (1.0 + e);
    }
    return fmin(1.0 - 1e-9, fmax(1e-9, q));
}

/* ln(1 + exp(-u)) as max(-u, 0) + log1p(exp(-|u|)): no overflow, no
   cancellation for small |u|. */
__device__ __forceinline__ double rc_softplus_neg(double u)
{
    const double a = fabs(u);
    return (u < 0.0 ? -u : 0.0) + log1p(exp(-a));
}

/* The logistic row terms are out of line: the branch is wave-uniform and
   rarely taken, and inlined f64 exp / log1p would cost the other metrics'
   kernels registers. */
__device__ __noinline__ double rc_softplus_neg_call(double u) { return rc_softplus_neg(u); }

/* The row loss of prediction p against the target slot y. */
__device__ __forceinline__ double rc_row_stat(double p, double y, int metric)
{
    if (metric == RC_LOSS_LOGISTIC) return rc_softplus_neg_call(y * p);
    return rc_row_rho(p - y, metric);
}

/* The row of the weighted least-squares step: s = sqrt(row weight), the
   residual r the row carries and the row loss rho.  Logistic: the Fisher
   weight w q (1 - q) with q = sigmoid(t p) and the residual -t / q, so that
   sum (s g)(s r) = -sum w t (1 - q) g is the loss gradient. */
struct RcRowFit {
    double s, r, rho;
};
__device__ __noinline__ RcRowFit rc_row_fit_logistic(double wb, double p, double y)
{
    RcRowFit o;
    const double u = y * p;
    const double q = rc_logistic_q(u);
    o.s = sqrt(wb * q * (1.0 - q));
    o.r = -y / q;
    o.rho = rc_softplus_neg(u);
    return o;
}
__device__ __forceinline__ RcRowFit rc_row_fit(double wb, double p, double y, int metric)
{
    RcRowFit o;
    if (metric == RC_LOSS_LOGISTIC) {
        o = rc_row_fit_logistic(wb, p, y);
    } else {
        o.r = p - y;
        o.s = sqrt(rc_row_weight(wb, o.r, metric));
        o.rho = rc_row_rho(o.r, metric);
    }
    return o;
}

/* ---- Wave-level helpers (wave64). */
__device__ __forceinline__ double rc_wave_sum_f64(double v)
{
    /* Fixed butterfly order: the result is independent of scheduling. */
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int rc_wave_or(int v)
{
    return __any(v) ? 1 : 0;
}

/* ---- Per-candidate reductions over work items (one lane).
   p[0] + p[stride] + ... + p[(n - 1) stride], added left to right (the value
   of the plain loop), with the loads issued eight at a time so that their
   latencies overlap instead of adding up.  n >= 0 (in the last batch an
   index past the end repeats the last element, which is loaded but not
   added). */
__device__ __forceinline__ double rc_ordered_sum(const double *__restrict__ p, size_t stride, int n)
{
    double acc = 0.0;
    for (int t0 = 0; t0 < n; t0 += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[(size_t)min(t0 + u, n - 1) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (t0 + u < n) acc += v[u];
    }
    return acc;
}

/* p[0] | p[1] | ... | p[n - 1], n >= 0, loads issued eight at a time. */
__device__ __forceinline__ int rc_batched_or(const int *__restrict__ p, int n)
{
    int acc = 0;
    for (int t0 = 0; t0 < n; t0 += 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[min(t0 + u, n - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) acc |= v[u];
    }
    return acc;
}

#endif
