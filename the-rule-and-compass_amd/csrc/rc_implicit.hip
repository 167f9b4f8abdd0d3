/*
 * Implicit relations on gfx950: f(x) = 0 scored and fitted by the signed
 * geometric distance d = f / ||grad_x f|| (implicit.rs:59-411, CPU only in
 * the reference; here a device arm behind rc_implicit_score_batch and
 * rc_implicit_fit_batch).
 *
 * Semantics per row, all in f64 as the reference:
 *   f        expr::eval with its domain guards (expr.rs:1054-1330),
 *   grad_x f the adjoint sweep over the fitter's value_of node values
 *            (fit.rs:1211-1365, 1453-1671) with the input gradients as
 *            output (fit.rs:3995 input_gradients),
 *   d        undefined when f, a gradient component, the norm or d is not
 *            finite, or the norm is not positive (signed_distance 68-96).
 *
 * Layout.  Work items are (candidate, 8192-row tile) with one wave each; a
 * lane walks its rows of the tile and evaluates the candidate's post-order
 * records (the rch_fit_plan form) with per-lane value and adjoint arrays.
 * Three kinds of pass:
 *   SCORE  sum w * |d| (metric 0) or w * |f| (metric 1) and sum w;
 *   RSS    sum (sqrt(w) d)^2 at one constant vector;
 *   JAC    the central-difference Jacobian of r = sqrt(w) d in the constants
 *          (2k + 1 distances per row, h = 1e-6 (1 + |c_j|), implicit.rs:
 *          320-345) and the augmented Gram [J | r]^T [J | r]: the rows of a
 *          64-row step are staged in LDS and every Gram entry is owned by one
 *          lane that adds the 64 products in row order.
 * Partials are per item and reduced on the host in item order, so a
 * candidate's result does not depend on the batch.  The LM control
 * (Gaussian elimination with partial pivoting, lambda x6 / /8, clamp to
 * +-1e6, implicit.rs:215-253, 277-409) runs on the host between batched
 * passes: one JAC and one RSS launch per iteration for the whole batch.
 */
#include "rc_device.h"
#include "rc_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int IM_TILE = 8192;
constexpr int IM_MAX_NODES = 96;
constexpr int IM_KMAX = 20;
constexpr int IM_STATS = (IM_KMAX + 1) * (IM_KMAX + 2) / 2; /* [J | r] lower triangle */
constexpr double IM_INV_SQRT_PI = 0.5641895835477563;

enum { IM_SCORE = 0, IM_RSS = 1, IM_JAC = 2 };

struct ImCand {
    uint32_t off; /* first record word */
    int m;        /* nodes */
    int k;        /* constants */
    int fast;     /* 1: runs the register-stack sweep (im_fast_kernel) */
};

/* The register-stack sweep's limits: values below the top of the stack
   (the post-order evaluation's depth less one) and input columns. */
constexpr int IM_SDEPTH = 8;
constexpr int IM_NVMAX = 4;
/* A stack file: two components of the 8 stack entries, one in elements
   0..7 and one in 8..15 (a 16-element vector is indexed with one register
   move per access; shorter ones are lowered to select chains). */
typedef double im_d16 __attribute__((ext_vector_type(16)));
typedef double im_d4 __attribute__((ext_vector_type(4)));

struct ImItem {
    int cand;
    int row0;
    int rows;
    int pad;
};

/* value_of (fit.rs:1211-1365) for the fit operator set. */
__device__ __forceinline__ double im_value(uint32_t op, double a, double b)
{
    switch (op) {
        case OP_ADD: return a + b;
        case OP_SUB: return a - b;
        case OP_MUL: return a * b;
        case OP_DIV: return fabs(b) < 1e-12 ? RC_F64_NAN : a / b;
        case OP_POW: return pow(a, b);
        case OP_NEG: return -a;
        case OP_EXP: return a > 700.0 ? __builtin_inf() : exp(a);
        case OP_LOG: return log(a);
        case OP_ABS: return fabs(a);
        case OP_SQRT: return sqrt(a);
        case OP_SIN: return sin(a);
        case OP_COS: return cos(a);
        case OP_TAN: return tan(a);
        case OP_SINH: return sinh(a);
        case OP_COSH: return cosh(a);
        case OP_TANH: return tanh(a);
        case OP_SIG: return 1.0 / (1.0 + exp(-a));
        case OP_GAUSS: return exp(-a * a);
        case OP_SQUARE: return a * a;
        case OP_CUBE: return a * a * a;
        case OP_ERF: return erf(a);
        case OP_ERFC: return erfc(a);
        default: return RC_F64_NAN;
    }
}

/* Whether expr::eval's guard fires where value_of does not (LOG of a
   non-positive, SQRT of a negative, a non-finite TAN, POW of a negative base
   to a non-integer power): then f comes from the guarded evaluation. */
__device__ __forceinline__ bool im_guard(uint32_t op, double a, double b, double v)
{
    switch (op) {
        case OP_LOG: return !(a > 0.0);
        case OP_SQRT: return !(a >= 0.0);
        case OP_TAN: return !isfinite(v);
        case OP_POW: return a < 0.0 && fabs(b - round(b)) > 1e-9;
        default: return false;
    }
}

__device__ __forceinline__ double im_guarded(uint32_t op, double a, double b)
{
    switch (op) {
        case OP_LOG: return a > 0.0 ? log(a) : RC_F64_NAN;
        case OP_SQRT: return a >= 0.0 ? sqrt(a) : RC_F64_NAN;
        case OP_TAN: { double t = tan(a); return isfinite(t) ? t : RC_F64_NAN; }
        case OP_POW: return (a < 0.0 && fabs(b - round(b)) > 1e-9) ? RC_F64_NAN : pow(a, b);
        default: return im_value(op, a, b);
    }
}

__device__ __forceinline__ bool im_is_bin(uint32_t op)
{
    return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV || op == OP_POW;
}

/* Signed distance (metric 0) or f itself (metric 1) of one row at constants
   c; false when the row is undefined.  v / adj are this lane's node arrays. */
__device__ __noinline__ bool im_distance(const uint32_t *__restrict__ r4, int m, const double *c,
                            const double *__restrict__ cols64, int n, int nv, int row,
                            int metric, double *v, double *adj, double &out)
{
    bool guard = false;
    for (int i = 0; i < m; ++i) {
        const uint32_t op = r4[4 * i], k0 = r4[4 * i + 1], k1 = r4[4 * i + 2], slot = r4[4 * i + 3];
        double x;
        if (op == OP_CONST) {
            x = c[slot];
        } else if (op == OP_VAR) {
            x = cols64[(size_t)slot * (size_t)n + (size_t)row];
        } else if (im_is_bin(op)) {
            const double a = v[k0], b = v[k1];
            x = im_value(op, a, b);
            guard = guard || im_guard(op, a, b, x);
        } else {
            const double a = v[k0];
            x = im_value(op, a, 0.0);
            guard = guard || im_guard(op, a, 0.0, x);
        }
        v[i] = x;
    }
    double fv = v[m - 1];
    if (guard) {
        /* expr::eval over the same records, in adj[] as scratch. */
        for (int i = 0; i < m; ++i) {
            const uint32_t op = r4[4 * i], k0 = r4[4 * i + 1], k1 = r4[4 * i + 2];
            if (op == OP_CONST || op == OP_VAR) adj[i] = v[i];
            else if (im_is_bin(op)) adj[i] = im_guarded(op, adj[k0], adj[k1]);
            else adj[i] = im_guarded(op, adj[k0], 0.0);
        }
        fv = adj[m - 1];
    }
    if (!isfinite(fv)) return false;
    if (metric == 1) {
        out = fv;
        return true;
    }
    /* Adjoint sweep (backward_impl) with the input gradients as output. */
    double g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < m; ++i) adj[i] = 0.0;
    adj[m - 1] = 1.0;
    for (int i = m - 1; i >= 0; --i) {
        const double a = adj[i];
        if (a == 0.0) continue;
        const uint32_t op = r4[4 * i], q0 = r4[4 * i + 1], q1 = r4[4 * i + 2], slot = r4[4 * i + 3];
        if (op == OP_CONST) continue;
        if (op == OP_VAR) {
            if ((int)slot < nv) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if ((int)slot == q) g[q] += a;
            }
            continue;
        }
        const double x = v[q0];
        switch (op) {
            case OP_NEG: adj[q0] += -a; break;
            case OP_EXP: adj[q0] += a * exp(x); break;
            case OP_LOG: adj[q0] += a / x; break;
            case OP_ABS: adj[q0] += a * (x >= 0.0 ? 1.0 : -1.0); break;
            case OP_SQRT: adj[q0] += a / (2.0 * sqrt(x)); break;
            case OP_SIN: adj[q0] += a * cos(x); break;
            case OP_COS: adj[q0] += -a * sin(x); break;
            case OP_TAN: { const double t = cos(x); adj[q0] += a / (t * t); } break;
            case OP_SINH: adj[q0] += a * cosh(x); break;
            case OP_COSH: adj[q0] += a * sinh(x); break;
            case OP_TANH: { const double t = tanh(x); adj[q0] += a * (1.0 - t * t); } break;
            case OP_SIG: {
                const double s = 1.0 / (1.0 + exp(-x));
                adj[q0] += a * s * (1.0 - s);
            } break;
            case OP_GAUSS: adj[q0] += a * (-2.0 * x) * exp(-x * x); break;
            case OP_SQUARE: adj[q0] += a * 2.0 * x; break;
            case OP_CUBE: adj[q0] += a * 3.0 * x * x; break;
            case OP_ERF: adj[q0] += a * 2.0 * IM_INV_SQRT_PI * exp(-x * x); break;
            case OP_ERFC: adj[q0] += -a * 2.0 * IM_INV_SQRT_PI * exp(-x * x); break;
            case OP_ADD: adj[q0] += a; adj[q1] += a; break;
            case OP_SUB: adj[q0] += a; adj[q1] += -a; break;
            case OP_MUL: adj[q0] += a * v[q1]; adj[q1] += a * x; break;
            case OP_DIV: {
                const double d = v[q1];
                adj[q0] += a / d;
                adj[q1] += -a * x / (d * d);
            } break;
            case OP_POW: {
                const double e = v[q1];
                if (x > 0.0) {
                    adj[q0] += a * e * pow(x, e - 1.0);
                    adj[q1] += a * pow(x, e) * log(x);
                }
            } break;
            default: break;
        }
    }
    double norm_sq = 0.0;
    for (int q = 0; q < 8; ++q) {
        if (q >= nv) break;
        if (!isfinite(g[q])) return false;
        norm_sq += g[q] * g[q];
    }
    const double norm = sqrt(norm_sq);
    if (!(isfinite(norm) && norm > 0.0)) return false;
    const double d = fv / norm;
    if (!isfinite(d)) return false;
    out = d;
    return true;
}

/* One pass over the work items.  cs holds, per candidate, the constant
   vector of this pass (kmax stride).  part receives per-item partials
   (pstride doubles), bad a per-item flag (1: a row with w > 0 is
   undefined, 2: a weight is negative or not finite). */
template <int MODE>
__global__ __launch_bounds__(64) void im_pass_kernel(
    const ImItem *__restrict__ items, int nitems, const ImCand *__restrict__ cands,
    const uint32_t *__restrict__ rec, const double *__restrict__ cs, int kmax,
    const double *__restrict__ cols64, const double *__restrict__ w64, int n, int nv, int metric,
    double *__restrict__ part, int pstride, int *__restrict__ bad, const int *__restrict__ run)
{
    const int it_i = blockIdx.x;
    if (it_i >= nitems) return;
    const ImItem it = items[it_i];
    if (run && !run[it.cand]) return;
    const ImCand d = cands[it.cand];
    if (d.fast) return;
    const int lane = threadIdx.x;
    const uint32_t *r4 = rec + d.off;
    const double *c = cs + (size_t)it.cand * (size_t)kmax;
    const int k = d.k;
    double v[IM_MAX_NODES], adj[IM_MAX_NODES];
    double cset[IM_KMAX];
    for (int j = 0; j < IM_KMAX; ++j) cset[j] = (j < k) ? c[j] : 0.0;
    __shared__ double J[64][IM_KMAX + 2];
    const int K1 = k + 1; /* [J | r] columns */
    const int nstats = K1 * (K1 + 1) / 2;
    double acc[4] = {0.0, 0.0, 0.0, 0.0}; /* owned Gram entries (JAC) */
    double s0 = 0.0, s1 = 0.0;
    int flag = 0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64) {
        const int row = base + lane;
        const bool live = row < row_end;
        const double w = (live && w64) ? w64[row] : 1.0;
        if (live && (!isfinite(w) || w < 0.0)) flag |= 2;
        const bool use = live && isfinite(w) && w > 0.0;
        if (MODE == IM_SCORE || MODE == IM_RSS) {
            if (use) {
                double dist;
                if (!im_distance(r4, d.m, cset, cols64, n, nv, row, MODE == IM_SCORE ? metric : 0,
                                 v, adj, dist)) {
                    flag |= 1;
                } else if (MODE == IM_SCORE) {
                    s0 += w * fabs(dist);
                    s1 += w;
                } else {
                    const double r = sqrt(w) * dist;
                    s0 += r * r;
                }
            }
        } else {
            double jrow[IM_KMAX + 1];
            for (int j = 0; j <= IM_KMAX; ++j) jrow[j] = 0.0;
            if (use) {
                const double sw = sqrt(w);
                double d0;
                bool ok = im_distance(r4, d.m, cset, cols64, n, nv, row, 0, v, adj, d0);
                jrow[k] = sw * d0;
                for (int j = 0; j < k && ok; ++j) {
                    const double cj = cset[j];
                    const double h = 1e-6 * (1.0 + fabs(cj));
                    double dp, dm;
                    cset[j] = cj + h;
                    ok = im_distance(r4, d.m, cset, cols64, n, nv, row, 0, v, adj, dp);
                    if (ok) {
                        cset[j] = cj - h;
                        ok = im_distance(r4, d.m, cset, cols64, n, nv, row, 0, v, adj, dm);
                    }
                    cset[j] = cj;
                    if (ok) jrow[j] = (sw * dp - sw * dm) / (2.0 * h);
                }
                if (!ok) flag |= 1;
                for (int j = 0; j <= k; ++j)
                    if (!isfinite(jrow[j])) flag |= 1;
            }
            for (int j = 0; j <= k; ++j) J[lane][j] = jrow[j];
            __syncthreads();
            for (int q = 0; q < 4; ++q) {
                const int s = lane + 64 * q;
                if (s >= nstats) break;
                /* packed lower triangle: s = a (a + 1) / 2 + b */
                int a = (int)((sqrt(8.0 * s + 1.0) - 1.0) * 0.5);
                while (a * (a + 1) / 2 > s) --a;
                while ((a + 1) * (a + 2) / 2 <= s) ++a;
                const int b = s - a * (a + 1) / 2;
                double t = acc[q];
                /* Rolled by 8: a full unroll hoists all 128 LDS reads and
                   spills. */
#pragma unroll 8
                for (int r = 0; r < 64; ++r) t += J[r][a] * J[r][b];
                acc[q] = t;
            }
            __syncthreads();
        }
    }
    double *p = part + (size_t)it_i * (size_t)pstride;
    int f = flag;
    for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off, 64);
    if (MODE == IM_JAC) {
        for (int q = 0; q < 4; ++q) {
            const int s = lane + 64 * q;
            if (s < nstats) p[s] = acc[q];
        }
    } else {
        const double t0 = rc_wave_sum_f64(s0);
        const double t1 = rc_wave_sum_f64(s1);
        if (lane == 0) {
            p[0] = t0;
            p[1] = t1;
        }
    }
    if (lane == 0) bad[it_i] = f;
}

/* ---- the register-stack sweep ------------------------------------------------

   The same per-row contract as im_distance for programs whose post-order
   evaluation keeps at most IM_SDEPTH values below the top of its stack, over
   at most IM_NVMAX input columns.  One forward sweep carries, per stack
   entry, the value_of value, the guarded (expr::eval) value and the
   tangents d/dx_q of the value_of value (forward-mode differentiation)
   instead of a value sweep, a guarded re-evaluation and an adjoint sweep
   over per-lane node arrays.  The stack lives in registers: the top in
   fixed registers, the entries below it in vectors indexed by the
   wave-uniform stack pointer.

   Every local partial is the adjoint sweep's (implicit.rs via the fitter's
   backward rules), applied to the tangents with the adjoint sweep's zero
   rule: a zero tangent or a zero partial contributes exactly 0, as a zero
   adjoint skips a node there.  The gradient therefore agrees with the
   adjoint sweep's up to the rounding of the products' association; where
   an exactly cancelling pair of adjoints (which the adjoint sweep skips)
   sits above an infinite partial, the two may differ, and both then flag
   the row in every case the tests hold. */
__device__ __forceinline__ double im_lm(double p, double t)
{
    return (p == 0.0 || t == 0.0) ? 0.0 : p * t;
}
__device__ __forceinline__ double im_ld(double t, double x) { return t == 0.0 ? 0.0 : t / x; }

/* A unary node of value_of argument a and guarded argument ga: its value,
   guarded value, whether a guard fires, and its local partial p (the
   tangent is t / p when div, else p * t, both under the zero rule).  Out
   of line, so the sweep's loop keeps few registers. */
struct ImUnary {
    double v, g, p;
    int div, guard;
};
__device__ __noinline__ ImUnary im_unary_all(uint32_t op, double a, double ga)
{
    ImUnary u;
    u.v = im_value(op, a, 0.0);
    /* The guarded evaluation runs the value's function on its own argument:
       when that is a itself (bit for bit) the result is the value, through
       the guarded operators' domain checks. */
    if (__double_as_longlong(ga) == __double_as_longlong(a)) {
        switch (op) {
            case OP_LOG: u.g = a > 0.0 ? u.v : RC_F64_NAN; break;
            case OP_SQRT: u.g = a >= 0.0 ? u.v : RC_F64_NAN; break;
            case OP_TAN: u.g = isfinite(u.v) ? u.v : RC_F64_NAN; break;
            default: u.g = u.v; break;
        }
    } else {
        u.g = im_guarded(op, ga, 0.0);
    }
    u.guard = im_guard(op, a, 0.0, u.v) ? 1 : 0;
    u.div = 0;
    /* A partial that is the value itself, or a multiple of it, reuses it (the
       same expression of the same argument). */
    switch (op) {
        case OP_NEG: u.p = -1.0; break;
        case OP_EXP: u.p = a > 700.0 ? exp(a) : u.v; break;
        case OP_LOG: u.p = a; u.div = 1; break;
        case OP_ABS: u.p = a >= 0.0 ? 1.0 : -1.0; break;
        case OP_SQRT: u.p = 2.0 * u.v; u.div = 1; break;
        case OP_SIN: u.p = cos(a); break;
        case OP_COS: u.p = -sin(a); break;
        case OP_TAN: { const double c = cos(a); u.p = c * c; u.div = 1; } break;
        case OP_SINH: u.p = cosh(a); break;
        case OP_COSH: u.p = sinh(a); break;
        case OP_TANH: { const double h = u.v; u.p = 1.0 - h * h; } break;
        case OP_SIG: { const double e = u.v; u.p = e * (1.0 - e); } break;
        case OP_GAUSS: u.p = (-2.0 * a) * u.v; break;
        case OP_SQUARE: u.p = 2.0 * a; break;
        case OP_CUBE: u.p = 3.0 * a * a; break;
        case OP_ERF: u.p = 2.0 * IM_INV_SQRT_PI * exp(-a * a); break;
        case OP_ERFC: u.p = -2.0 * IM_INV_SQRT_PI * exp(-a * a); break;
        default: u.p = 0.0; break;
    }
    return u;
}

/* A POW node: value, guarded value, guard, and the partials of both
   arguments (zero unless the base is positive). */
struct ImPow {
    double v, g, pa, pb;
    int guard;
};
__device__ __noinline__ ImPow im_pow_all(double a, double b, double ga, double gb)
{
    ImPow w;
    w.v = im_value(OP_POW, a, b);
    w.g = im_guarded(OP_POW, ga, gb);
    w.guard = im_guard(OP_POW, a, b, w.v) ? 1 : 0;
    if (a > 0.0) {
        w.pa = b * pow(a, b - 1.0);
        w.pb = pow(a, b) * log(a);
    } else {
        w.pa = 0.0;
        w.pb = 0.0;
    }
    return w;
}

/* The signed distance (metric 0) or f (metric 1) of one row at the
   constants c, with constant slot pj replaced by cpj (pj < 0: none); x
   holds the row's NV inputs.  False when the row is undefined. */
template <int NV>
__device__ __forceinline__ bool im_fast_distance(const uint32_t *__restrict__ r4, int m,
                                                 const double *__restrict__ c, int pj, double cpj,
                                                 const double (&x)[IM_NVMAX], int metric,
                                                 double &out)
{
    im_d16 FV, F01, F23; /* value | guarded value, t0 | t1, t2 | t3 */
    double tv = 0.0, tg = 0.0, t[IM_NVMAX] = {0.0, 0.0, 0.0, 0.0};
    int sp = -1; /* entries below the top; -1: empty */
    bool guard = false;
    for (int i = 0; i < m; ++i) {
        const uint32_t op = r4[4 * i], slot = r4[4 * i + 3];
        if (op == OP_CONST || op == OP_VAR) {
            if (sp >= 0) {
                FV[sp] = tv;
                FV[sp + 8] = tg;
                F01[sp] = t[0];
                if (NV > 1) F01[sp + 8] = t[1];
                if (NV > 2) F23[sp] = t[2];
                if (NV > 3) F23[sp + 8] = t[3];
            }
            ++sp;
            if (op == OP_CONST) {
                tv = ((int)slot == pj) ? cpj : c[slot];
#pragma unroll
                for (int q = 0; q < IM_NVMAX; ++q) t[q] = 0.0;
            } else {
                tv = slot == 0 ? x[0] : slot == 1 ? x[1] : slot == 2 ? x[2] : x[3];
#pragma unroll
                for (int q = 0; q < IM_NVMAX; ++q) t[q] = ((int)slot == q) ? 1.0 : 0.0;
            }
            tg = tv;
        } else if (im_is_bin(op)) {
            --sp;
            const double va = FV[sp], ga = FV[sp + 8];
            double ta[IM_NVMAX];
            ta[0] = F01[sp];
            ta[1] = NV > 1 ? F01[sp + 8] : 0.0;
            ta[2] = NV > 2 ? F23[sp] : 0.0;
            ta[3] = NV > 3 ? F23[sp + 8] : 0.0;
            const double vb = tv, gb = tg;
            if (op == OP_POW) {
                const ImPow w = im_pow_all(va, vb, ga, gb);
                guard = guard || w.guard;
                tg = w.g;
                tv = w.v;
#pragma unroll
                for (int q = 0; q < NV; ++q) t[q] = im_lm(w.pa, ta[q]) + im_lm(w.pb, t[q]);
            } else {
                /* ADD, SUB, MUL, DIV: no guard fires; the guarded value is
                   value_of's on the guarded arguments. */
                auto arith = [op](double l, double r) {
                    if (op == OP_ADD) return l + r;
                    if (op == OP_SUB) return l - r;
                    if (op == OP_MUL) return l * r;
                    return fabs(r) < 1e-12 ? RC_F64_NAN : l / r; /* value_of's DIV */
                };
                tv = arith(va, vb);
                tg = arith(ga, gb);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const double a_ = ta[q], b_ = t[q];
                    double r;
                    if (op == OP_ADD) r = a_ + b_;
                    else if (op == OP_SUB) r = a_ - b_;
                    else if (op == OP_MUL) r = im_lm(vb, a_) + im_lm(va, b_);
                    else r = im_ld(a_, vb) + (b_ == 0.0 ? 0.0 : -(va * b_) / (vb * vb));
                    t[q] = r;
                }
            }
        } else {
            const ImUnary u = im_unary_all(op, tv, tg);
            guard = guard || u.guard;
            tg = u.g;
            tv = u.v;
#pragma unroll
            for (int q = 0; q < NV; ++q) t[q] = u.div ? im_ld(t[q], u.p) : im_lm(u.p, t[q]);
        }
    }
    const double fv = guard ? tg : tv;
    if (!isfinite(fv)) return false;
    if (metric == 1) {
        out = fv;
        return true;
    }
    double norm_sq = 0.0;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        if (!isfinite(t[q])) return false;
        norm_sq += t[q] * t[q];
    }
    const double norm = sqrt(norm_sq);
    if (!(isfinite(norm) && norm > 0.0)) return false;
    const double dd = fv / norm;
    if (!isfinite(dd)) return false;
    out = dd;
    return true;
}

/* im_pass_kernel on the register-stack sweep for the candidates marked
   fast (the general kernel skips them, this one skips the others); NV is
   the row set's input count. */
template <int MODE, int NV>
__global__ __launch_bounds__(64) void im_fast_kernel(
    const ImItem *__restrict__ items, int nitems, const ImCand *__restrict__ cands,
    const uint32_t *__restrict__ rec, const double *__restrict__ cs, int kmax,
    const double *__restrict__ cols64, const double *__restrict__ w64, int n, int metric,
    double *__restrict__ part, int pstride, int *__restrict__ bad, const int *__restrict__ run)
{
    const int it_i = blockIdx.x;
    if (it_i >= nitems) return;
    const ImItem it = items[it_i];
    if (run && !run[it.cand]) return;
    const ImCand d = cands[it.cand];
    if (!d.fast) return;
    const int lane = threadIdx.x;
    const uint32_t *r4 = rec + d.off;
    const double *c = cs + (size_t)it.cand * (size_t)kmax;
    const int k = d.k;
    __shared__ double J[64][IM_KMAX + 2];
    const int K1 = k + 1; /* [J | r] columns */
    /* The Gram of [J | r] on FP64 MFMA: block (0, 0) of columns 0..15 and,
       past 16 columns, blocks (1, 0) and (1, 1). */
    im_d4 acc00 = {0.0, 0.0, 0.0, 0.0}, acc10 = acc00, acc11 = acc00;
    if (MODE == IM_JAC) {
        for (int q = lane; q < 64 * (IM_KMAX + 2); q += 64) (&J[0][0])[q] = 0.0;
        __syncthreads();
    }
    double s0 = 0.0, s1 = 0.0;
    int flag = 0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64) {
        const int row = base + lane;
        const bool live = row < row_end;
        const double w = (live && w64) ? w64[row] : 1.0;
        if (live && (!isfinite(w) || w < 0.0)) flag |= 2;
        const bool use = live && isfinite(w) && w > 0.0;
        double x[IM_NVMAX] = {0.0, 0.0, 0.0, 0.0};
        if (use) {
#pragma unroll
            for (int q = 0; q < NV; ++q) x[q] = cols64[(size_t)q * (size_t)n + (size_t)row];
        }
        if (MODE == IM_SCORE || MODE == IM_RSS) {
            if (use) {
                double dist;
                if (!im_fast_distance<NV>(r4, d.m, c, -1, 0.0, x, MODE == IM_SCORE ? metric : 0,
                                          dist)) {
                    flag |= 1;
                } else if (MODE == IM_SCORE) {
                    s0 += w * fabs(dist);
                    s1 += w;
                } else {
                    const double r = sqrt(w) * dist;
                    s0 += r * r;
                }
            }
        } else {
            for (int j = 0; j <= k; ++j) J[lane][j] = 0.0;
            if (use) {
                /* The 2k + 1 distances of the central difference in one
                   loop (one copy of the sweep): e = 0 at c, then c_j + h
                   and c_j - h for each j.  A row stops at its first
                   undefined distance, as the node-array kernel's does. */
                const double sw = sqrt(w);
                bool ok = true, fin = true;
                double dp = 0.0;
                for (int e = 0; e <= 2 * k && ok; ++e) {
                    const int j = (e - 1) >> 1;
                    double cp = 0.0, h = 0.0;
                    if (e > 0) {
                        const double cj = c[j];
                        h = 1e-6 * (1.0 + fabs(cj));
                        cp = (e & 1) ? cj + h : cj - h;
                    }
                    double dist = 0.0;
                    ok = im_fast_distance<NV>(r4, d.m, c, e > 0 ? j : -1, cp, x, 0, dist);
                    if (!ok) break;
                    if (e == 0) {
                        J[lane][k] = sw * dist;
                        fin = isfinite(J[lane][k]);
                    } else if (e & 1) {
                        dp = dist;
                    } else {
                        const double g = (sw * dp - sw * dist) / (2.0 * h);
                        J[lane][j] = g;
                        fin = fin && isfinite(g);
                    }
                }
                if (!ok || !fin) flag |= 1;
            }
            __syncthreads();
            {
                /* Operand of lane (col, sub) for rows 4 t + sub: A and B of
                   the MFMA are the same block of J. */
                const int col = lane & 15, sub = lane >> 4;
                const int col1 = (16 + col < IM_KMAX + 2) ? 16 + col : IM_KMAX + 1;
#pragma unroll 4
                for (int t4 = 0; t4 < 16; ++t4) {
                    const double a0 = J[4 * t4 + sub][col];
                    acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc00, 0, 0, 0);
                    if (K1 > 16) {
                        const double a1 = J[4 * t4 + sub][col1];
                        acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a0, acc10, 0, 0, 0);
                        acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, acc11, 0, 0, 0);
                    }
                }
            }
            __syncthreads();
        }
    }
    double *p = part + (size_t)it_i * (size_t)pstride;
    int f = flag;
    for (int off = 32; off > 0; off >>= 1) f |= __shfl_xor(f, off, 64);
    if (MODE == IM_JAC) {
        /* FP64 MFMA output: lane l, register r holds D[l/16 + 4 r][l % 16]
           of each 16x16 block; p is the packed lower triangle
           a (a + 1) / 2 + b. */
        const int jj = lane & 15;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int ii = (lane >> 4) + 4 * r;
            if (ii >= jj && ii < K1) p[ii * (ii + 1) / 2 + jj] = acc00[r];
            const int i1 = 16 + ii;
            if (i1 < K1) p[i1 * (i1 + 1) / 2 + jj] = acc10[r];
            if (ii >= jj && i1 < K1) p[i1 * (i1 + 1) / 2 + 16 + jj] = acc11[r];
        }
    } else {
        const double t0 = rc_wave_sum_f64(s0);
        const double t1 = rc_wave_sum_f64(s1);
        if (lane == 0) {
            p[0] = t0;
            p[1] = t1;
        }
    }
    if (lane == 0) bad[it_i] = f;
}

/* ---- the LM control of the implicit fit, on the device ------------------------

   One wave per candidate; the state lives in HBM between launches.  The
   launch sequence of a fit is: INIT, an RSS pass at the start constants,
   BASE, then per LM iteration a JAC pass, AFTER_JAC (the damped solve and
   the trial), an RSS pass and AFTER_RSS (accept / reject), and FINAL.  The
   passes run only the items of candidates whose run flag the previous
   control step set, so no host round trip sits between iterations.
   Semantics: implicit.rs:277-409 (lambda 1e-3, x6 on a rejected or failed
   step, /8 (floor 1e-12) on an accepted one, retirement above 1e9, steps
   clamped to +-1e6, best-so-far constants). */
struct ImState {
    double c[IM_KMAX];
    double trial[IM_KMAX];
    double base, best, lambda;
    int active, want, pad0, pad1;
};

enum { IC_INIT = 0, IC_BASE = 1, IC_AFTER_JAC = 2, IC_AFTER_RSS = 3, IC_FINAL = 4 };

/* implicit.rs:215-253 on the k x (k + 1) augmented matrix M (row stride
   IM_KMAX + 1): partial pivoting on the first largest magnitude, a pivot
   below 1e-300 fails, then back substitution into d.  Every product and
   sum is rounded on its own, as in the reference's solve. */
__device__ bool im_solve_dev(double *M, int k, double *d)
{
    /* Separate roundings, as the host solve: no fused multiply-adds. */
#pragma clang fp contract(off)
    const int S = IM_KMAX + 1;
    for (int col = 0; col < k; ++col) {
        int piv = col;
        for (int r = col + 1; r < k; ++r)
            if (fabs(M[r * S + col]) > fabs(M[piv * S + col])) piv = r;
        if (fabs(M[piv * S + col]) < 1e-300) return false;
        if (piv != col)
            for (int q = 0; q <= k; ++q) {
                const double t = M[col * S + q];
                M[col * S + q] = M[piv * S + q];
                M[piv * S + q] = t;
            }
        for (int r = col + 1; r < k; ++r) {
            const double factor = M[r * S + col] / M[col * S + col];
            if (factor != 0.0)
                for (int q = col; q <= k; ++q) M[r * S + q] -= factor * M[col * S + q];
        }
    }
    for (int i = k - 1; i >= 0; --i) {
        double acc = M[i * S + k];
        for (int j = i + 1; j < k; ++j) acc -= M[i * S + j] * d[j];
        d[i] = acc / M[i * S + i];
    }
    return true;
}

__global__ __launch_bounds__(64) void im_ctrl_kernel(
    int op, int np, ImState *__restrict__ S, const ImCand *__restrict__ cands,
    const int *__restrict__ pre, const double *__restrict__ c0, int kmax, int iters,
    double *__restrict__ cs, int *__restrict__ run, const int *__restrict__ item0,
    const int *__restrict__ nitem, const double *__restrict__ part, const int *__restrict__ bad,
    double *__restrict__ outc, double *__restrict__ outsse, int *__restrict__ outstatus)
{
#pragma clang fp contract(off)
    const int cand = blockIdx.x;
    if (cand >= np) return;
    const int lane = threadIdx.x;
    __shared__ ImState st;
    __shared__ double G[IM_STATS];
    __shared__ double M[IM_KMAX * (IM_KMAX + 1)];
    constexpr int SW = (int)(sizeof(ImState) / sizeof(int));
    static_assert(sizeof(ImState) % sizeof(int) == 0, "ImState is copied as ints");
    for (int i = lane; i < SW; i += 64) ((int *)&st)[i] = ((const int *)(S + cand))[i];
    __syncthreads();
    const int k = cands[cand].k;
    const bool host_done = pre[cand] >= 0;
    const bool ran = run[cand] != 0;
    const int i0 = item0[cand], ni = nitem[cand];
    if (op == IC_AFTER_JAC && !host_done && st.active && ran) {
        const int nst = (k + 1) * (k + 2) / 2;
        for (int t = lane; t < nst; t += 64)
            G[t] = rc_ordered_sum(part + (size_t)i0 * IM_STATS + t, IM_STATS, ni);
    }
    __syncthreads();
    if (lane == 0) {
        double *cc = cs + (size_t)cand * (size_t)kmax;
        auto flag_sum = [&](double &sum) {
            sum = rc_ordered_sum(part + (size_t)i0 * IM_STATS, IM_STATS, ni);
            return rc_batched_or(bad + i0, ni);
        };
        int next_run = 0;
        switch (op) {
            case IC_INIT: {
                for (int j = 0; j < IM_KMAX; ++j) st.c[j] = j < kmax ? c0[(size_t)cand * kmax + j] : 0.0;
                st.base = st.best = __builtin_inf();
                st.lambda = 1e-3;
                st.active = 0;
                st.want = 0;
                for (int j = 0; j < kmax; ++j) cc[j] = st.c[j];
                next_run = !host_done;
                break;
            }
            case IC_BASE: {
                if (host_done || !ran) break;
                double sum;
                const int f = flag_sum(sum);
                const bool ok = f == 0 && isfinite(sum);
                st.base = st.best = ok ? sum : __builtin_inf();
                st.active = ok && k > 0 && iters > 0;
                next_run = st.active;
                break;
            }
            case IC_AFTER_JAC: {
                st.want = 0;
                if (host_done || !st.active || !ran) break;
                int f = 0;
                for (int u = 0; u < ni; ++u) f |= bad[i0 + u];
                if (f != 0) {
                    st.active = 0;
                    break;
                }
                const int Sx = IM_KMAX + 1;
                bool fin = true;
                for (int a = 0; a < k; ++a) {
                    for (int b = 0; b <= a; ++b) {
                        const double v = G[a * (a + 1) / 2 + b];
                        M[a * Sx + b] = v;
                        M[b * Sx + a] = v;
                        fin = fin && isfinite(v);
                    }
                    M[a * Sx + k] = -G[k * (k + 1) / 2 + a];
                    fin = fin && isfinite(M[a * Sx + k]);
                }
                if (!fin) {
                    st.active = 0;
                    break;
                }
                for (int a = 0; a < k; ++a) M[a * Sx + a] = M[a * Sx + a] * (1.0 + st.lambda) + 1e-12;
                double step[IM_KMAX];
                if (!im_solve_dev(M, k, step)) {
                    st.lambda *= 6.0;
                    if (st.lambda > 1e9) st.active = 0;
                    break;
                }
                for (int j = 0; j < kmax; ++j) {
                    const double t = j < k ? fmin(1e6, fmax(-1e6, st.c[j] + step[j])) : st.c[j];
                    st.trial[j] = t;
                    cc[j] = t;
                }
                st.want = 1;
                next_run = 1;
                break;
            }
            case IC_AFTER_RSS: {
                if (host_done || !st.want || !ran) {
                    next_run = !host_done && st.active;
                    break;
                }
                double sum;
                const int f = flag_sum(sum);
                const double t = f == 0 ? sum : __builtin_inf();
                if (f == 0 && t < st.best) {
                    st.best = t;
                    for (int j = 0; j < k; ++j) st.c[j] = st.trial[j];
                    st.lambda = fmax(st.lambda / 8.0, 1e-12);
                    if (st.best == 0.0) st.active = 0;
                } else {
                    st.lambda *= 6.0;
                    if (st.lambda > 1e9) st.active = 0;
                }
                st.want = 0;
                for (int j = 0; j < k; ++j) cc[j] = st.c[j];
                next_run = st.active;
                break;
            }
            case IC_FINAL: {
                double *oc = outc + (size_t)cand * kmax;
                const double *sc = c0 + (size_t)cand * kmax;
                if (host_done) {
                    for (int j = 0; j < kmax; ++j) oc[j] = sc[j];
                    outsse[cand] = __builtin_nan("");
                    outstatus[cand] = 1;
                } else if (!isfinite(st.base)) {
                    for (int j = 0; j < kmax; ++j) oc[j] = sc[j];
                    outsse[cand] = __builtin_inf();
                    outstatus[cand] = 7;
                } else {
                    const bool moved = st.best < st.base;
                    for (int j = 0; j < kmax; ++j) oc[j] = (moved && j < k) ? st.c[j] : sc[j];
                    outsse[cand] = st.best;
                    outstatus[cand] = 0;
                }
                break;
            }
            default: break;
        }
        run[cand] = next_run;
    }
    __syncthreads();
    for (int i = lane; i < SW; i += 64) ((int *)(S + cand))[i] = ((const int *)&st)[i];
}

/* Host side: validated candidates, resident buffers, pass launcher. */
struct ImBatch {
    RcCtx *c = nullptr;
    int nprog = 0, kmax = 0;
    std::vector<ImCand> cands;
    std::vector<int> pre; /* -1 device, else a host status */
    std::vector<uint32_t> rec;
    uint32_t *d_rec = nullptr;
    ImCand *d_cand = nullptr;
    ImItem *d_items = nullptr;
    double *d_cs = nullptr, *d_part = nullptr;
    int *d_bad = nullptr;
    std::vector<ImItem> h_items;
    std::vector<double> h_part;
    std::vector<int> h_bad;
    float ms = 0.0f;
    bool any_fast = false, any_general = false;

    /* One pass over ni items of mode MODE on stream s: the register-stack
       kernel for the fast candidates, the node-array kernel for the rest. */
    template <int MODE>
    void launch(hipStream_t s, int ni, int metric, const int *run)
    {
        if (ni <= 0) return;
        RcRows &R = c->sets[0];
        if (any_fast) {
            switch (R.v) {
                case 1: im_fast_kernel<MODE, 1><<<ni, 64, 0, s>>>(d_items, ni, d_cand, d_rec, d_cs, kmax, R.cols64, R.w64, R.n, metric, d_part, IM_STATS, d_bad, run); break;
                case 2: im_fast_kernel<MODE, 2><<<ni, 64, 0, s>>>(d_items, ni, d_cand, d_rec, d_cs, kmax, R.cols64, R.w64, R.n, metric, d_part, IM_STATS, d_bad, run); break;
                case 3: im_fast_kernel<MODE, 3><<<ni, 64, 0, s>>>(d_items, ni, d_cand, d_rec, d_cs, kmax, R.cols64, R.w64, R.n, metric, d_part, IM_STATS, d_bad, run); break;
                default: im_fast_kernel<MODE, 4><<<ni, 64, 0, s>>>(d_items, ni, d_cand, d_rec, d_cs, kmax, R.cols64, R.w64, R.n, metric, d_part, IM_STATS, d_bad, run); break;
            }
        }
        if (any_general)
            im_pass_kernel<MODE><<<ni, 64, 0, s>>>(d_items, ni, d_cand, d_rec, d_cs, kmax, R.cols64,
                                                   R.w64, R.n, R.v, metric, d_part, IM_STATS, d_bad,
                                                   run);
    }

    ~ImBatch()
    {
        if (d_rec) (void)hipFree(d_rec);
        if (d_cand) (void)hipFree(d_cand);
        if (d_items) (void)hipFree(d_items);
        if (d_cs) (void)hipFree(d_cs);
        if (d_part) (void)hipFree(d_part);
        if (d_bad) (void)hipFree(d_bad);
    }

    /* Post-order validation as the fit path's (canonical children, supported
       operators, slots in range). */
    int setup(RcCtx *ctx, const unsigned *prog, const unsigned *off, const int *pcount, int np,
              int km)
    {
        c = ctx;
        nprog = np;
        kmax = km;
        RcRows &R = c->sets[0];
        cands.assign((size_t)np, ImCand{0, 0, 0, 0});
        pre.assign((size_t)np, -1);
        for (int i = 0; i < np; ++i) {
            const int words = (int)(off[i + 1] - off[i]);
            const int m = words / 4;
            const unsigned *w = prog + off[i];
            const int p = pcount[i];
            bool ok = words % 4 == 0 && m >= 1 && m <= IM_MAX_NODES && p >= 0 && p <= kmax &&
                      p <= IM_KMAX;
            std::vector<int> st;
            size_t depth = 0;
            for (int q = 0; ok && q < m; ++q) {
                const uint32_t op = w[4 * q], k0 = w[4 * q + 1], k1 = w[4 * q + 2], slot = w[4 * q + 3];
                if (op == OP_CONST) {
                    ok = (int)slot < p;
                    st.push_back(q);
                } else if (op == OP_VAR) {
                    ok = (int)slot < R.v && slot < 8u;
                    st.push_back(q);
                } else if (op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                           op == OP_POW) {
                    ok = st.size() >= 2 && st.back() == (int)k1 && st[st.size() - 2] == (int)k0;
                    if (ok) { st.pop_back(); st.pop_back(); st.push_back(q); }
                } else {
                    const bool known = op == OP_SIN || op == OP_COS || op == OP_TAN || op == OP_EXP ||
                                       op == OP_LOG || op == OP_SQRT || op == OP_ABS ||
                                       op == OP_TANH || op == OP_SIG || op == OP_SQUARE ||
                                       op == OP_CUBE || op == OP_NEG || op == OP_GAUSS ||
                                       op == OP_SINH || op == OP_COSH || op == OP_ERF ||
                                       op == OP_ERFC;
                    ok = known && !st.empty() && st.back() == (int)k0;
                    if (ok) { st.pop_back(); st.push_back(q); }
                }
                depth = std::max(depth, st.size());
            }
            ok = ok && st.size() == 1 && st[0] == m - 1;
            if (!ok) {
                pre[i] = 1;
                continue;
            }
            cands[i].off = (uint32_t)rec.size();
            cands[i].m = m;
            cands[i].k = p;
            cands[i].fast = (depth <= (size_t)IM_SDEPTH + 1 && R.v >= 1 && R.v <= IM_NVMAX) ? 1 : 0;
            (cands[i].fast ? any_fast : any_general) = true;
            rec.insert(rec.end(), w, w + words);
        }
        if (rec.empty()) rec.push_back(0);
        for (int i = 0; i < np; ++i) {
            if (pre[i] >= 0) continue;
            for (int r0 = 0; r0 < R.n; r0 += IM_TILE)
                h_items.push_back(ImItem{i, r0, std::min(IM_TILE, R.n - r0), 0});
        }
        const size_t ni = std::max<size_t>(1, h_items.size());
        if (hipMalloc((void **)&d_rec, rec.size() * 4) != hipSuccess ||
            hipMalloc((void **)&d_cand, sizeof(ImCand) * (size_t)np) != hipSuccess ||
            hipMalloc((void **)&d_items, sizeof(ImItem) * ni) != hipSuccess ||
            hipMalloc((void **)&d_cs, sizeof(double) * (size_t)np * (size_t)kmax) != hipSuccess ||
            hipMalloc((void **)&d_part, sizeof(double) * ni * IM_STATS) != hipSuccess ||
            hipMalloc((void **)&d_bad, sizeof(int) * ni) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        hipStream_t s = c->s_fit;
        if (hipMemcpyAsync(d_rec, rec.data(), rec.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d_cand, cands.data(), sizeof(ImCand) * (size_t)np, hipMemcpyHostToDevice, s) != hipSuccess ||
            (!h_items.empty() && hipMemcpyAsync(d_items, h_items.data(), sizeof(ImItem) * h_items.size(),
                                                hipMemcpyHostToDevice, s) != hipSuccess))
            return -1;
        h_part.resize(ni * IM_STATS);
        h_bad.resize(ni);
        return 0;
    }

    /* Runs one pass over the items of the candidates with run[i] != 0 at
       constants cs (np x kmax); fills per-candidate sums: out[i][0..] and the
       OR of the item flags. */
    int pass(int mode, int metric, const std::vector<double> &cs, const std::vector<char> &run,
             std::vector<std::vector<double>> &sums, std::vector<int> &flags)
    {
        std::vector<ImItem> sel;
        std::vector<int> from;
        for (size_t q = 0; q < h_items.size(); ++q)
            if (run[(size_t)h_items[q].cand]) {
                sel.push_back(h_items[q]);
                from.push_back((int)q);
            }
        sums.assign((size_t)nprog, std::vector<double>());
        flags.assign((size_t)nprog, 0);
        if (sel.empty()) return 0;
        hipStream_t s = c->s_fit;
        if (hipMemcpyAsync(d_items, sel.data(), sizeof(ImItem) * sel.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(d_cs, cs.data(), sizeof(double) * cs.size(), hipMemcpyHostToDevice, s) != hipSuccess)
            return -1;
        const int ni = (int)sel.size();
        (void)hipEventRecord(c->t0, s);
        if (mode == IM_SCORE) launch<IM_SCORE>(s, ni, metric, nullptr);
        else if (mode == IM_RSS) launch<IM_RSS>(s, ni, metric, nullptr);
        else launch<IM_JAC>(s, ni, metric, nullptr);
        (void)hipEventRecord(c->t1, s);
        if (hipGetLastError() != hipSuccess) return -1;
        if (hipMemcpyAsync(h_part.data(), d_part, sizeof(double) * (size_t)ni * IM_STATS,
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(h_bad.data(), d_bad, sizeof(int) * (size_t)ni, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return -1;
        float e = 0.0f;
        (void)hipEventElapsedTime(&e, c->t0, c->t1);
        ms += e;
        /* Items of one candidate are consecutive and in row order. */
        for (int q = 0; q < ni; ++q) {
            const int cand = sel[(size_t)q].cand;
            const int K1 = cands[(size_t)cand].k + 1;
            const int nst = (mode == IM_JAC) ? K1 * (K1 + 1) / 2 : 2;
            std::vector<double> &t = sums[(size_t)cand];
            if (t.empty()) t.assign((size_t)nst, 0.0);
            for (int j = 0; j < nst; ++j) t[(size_t)j] += h_part[(size_t)q * IM_STATS + j];
            flags[(size_t)cand] |= h_bad[(size_t)q];
        }
        return 0;
    }
};

constexpr double IM_INF = __builtin_inf();

} /* namespace */

int rc_rows_ensure_f64(RcCtx *c, RcRows *r);

extern "C" {

float rc_implicit_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                              const int *pcount, int nprog, int kmax, const double *c, int metric,
                              double *out, int *status)
{
    RcCtx *cx = (RcCtx *)ctx;
    if (!cx || !prog || !off || !pcount || !c || !out || !status || nprog <= 0 || kmax <= 0 ||
        kmax > IM_KMAX || (metric != 0 && metric != 1))
        return -1.0f;
    if (rc_rows_ensure_f64(cx, &cx->sets[0]) != 0) return -1.0f;
    ImBatch B;
    if (B.setup(cx, prog, off, pcount, nprog, kmax) != 0) return -1.0f;
    std::vector<double> cs(c, c + (size_t)nprog * (size_t)kmax);
    std::vector<char> run((size_t)nprog, 1);
    for (int i = 0; i < nprog; ++i) run[i] = B.pre[i] < 0;
    std::vector<std::vector<double>> sums;
    std::vector<int> flags;
    if (B.pass(IM_SCORE, metric, cs, run, sums, flags) != 0) return -1.0f;
    for (int i = 0; i < nprog; ++i) {
        if (B.pre[i] >= 0) {
            out[i] = RC_F64_NAN;
            status[i] = 1;
            continue;
        }
        const std::vector<double> &t = sums[(size_t)i];
        const bool fin = !t.empty() && flags[i] == 0 && isfinite(t[1]) && t[1] > 0.0;
        out[i] = fin ? t[0] / t[1] : IM_INF;
        status[i] = fin ? 0 : 2;
    }
    return B.ms;
}

float rc_implicit_fit_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                            const int *pcount, int nprog, int kmax, const double *c0, int iters,
                            double *outc, double *outsse, int *outstatus)
{
    RcCtx *cx = (RcCtx *)ctx;
    if (!cx || !prog || !off || !pcount || !c0 || !outc || !outsse || !outstatus || nprog <= 0 ||
        kmax <= 0 || kmax > IM_KMAX || iters < 0)
        return -1.0f;
    if (rc_rows_ensure_f64(cx, &cx->sets[0]) != 0) return -1.0f;
    ImBatch B;
    if (B.setup(cx, prog, off, pcount, nprog, kmax) != 0) return -1.0f;
    const size_t P = (size_t)nprog, KM = (size_t)kmax;
    /* Each candidate's items are consecutive in row order. */
    std::vector<int> item0(P, 0), nitem(P, 0);
    for (size_t q = 0; q < B.h_items.size(); ++q) {
        const int cd = B.h_items[q].cand;
        if (nitem[(size_t)cd] == 0) item0[(size_t)cd] = (int)q;
        nitem[(size_t)cd] += 1;
    }
    ImState *d_state = nullptr;
    int *d_run = nullptr, *d_pre = nullptr, *d_i0 = nullptr, *d_ni = nullptr, *d_st = nullptr;
    double *d_c0 = nullptr, *d_outc = nullptr, *d_sse = nullptr;
    auto release = [&]() {
        if (d_state) (void)hipFree(d_state);
        if (d_run) (void)hipFree(d_run);
        if (d_pre) (void)hipFree(d_pre);
        if (d_i0) (void)hipFree(d_i0);
        if (d_ni) (void)hipFree(d_ni);
        if (d_st) (void)hipFree(d_st);
        if (d_c0) (void)hipFree(d_c0);
        if (d_outc) (void)hipFree(d_outc);
        if (d_sse) (void)hipFree(d_sse);
    };
    if (hipMalloc((void **)&d_state, sizeof(ImState) * P) != hipSuccess ||
        hipMalloc((void **)&d_run, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void **)&d_pre, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void **)&d_i0, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void **)&d_ni, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void **)&d_st, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void **)&d_c0, sizeof(double) * P * KM) != hipSuccess ||
        hipMalloc((void **)&d_outc, sizeof(double) * P * KM) != hipSuccess ||
        hipMalloc((void **)&d_sse, sizeof(double) * P) != hipSuccess) {
        (void)hipGetLastError();
        release();
        return -1.0f;
    }
    hipStream_t s = cx->s_fit;
    const int ni = (int)B.h_items.size();
    bool ok = hipMemcpyAsync(d_pre, B.pre.data(), sizeof(int) * P, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_i0, item0.data(), sizeof(int) * P, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_ni, nitem.data(), sizeof(int) * P, hipMemcpyHostToDevice, s) == hipSuccess &&
              hipMemcpyAsync(d_c0, c0, sizeof(double) * P * KM, hipMemcpyHostToDevice, s) == hipSuccess;
    auto ctrl = [&](int op) {
        im_ctrl_kernel<<<nprog, 64, 0, s>>>(op, nprog, d_state, B.d_cand, d_pre, d_c0, kmax, iters,
                                            B.d_cs, d_run, d_i0, d_ni, B.d_part, B.d_bad, d_outc,
                                            d_sse, d_st);
    };
    auto pass = [&](int mode) {
        if (mode == IM_RSS) B.launch<IM_RSS>(s, ni, 0, d_run);
        else B.launch<IM_JAC>(s, ni, 0, d_run);
    };
    if (ok) {
        (void)hipEventRecord(cx->t0, s);
        ctrl(IC_INIT);
        pass(IM_RSS);
        ctrl(IC_BASE);
        for (int it = 0; it < std::max(iters, 1); ++it) {
            pass(IM_JAC);
            ctrl(IC_AFTER_JAC);
            pass(IM_RSS);
            ctrl(IC_AFTER_RSS);
        }
        ctrl(IC_FINAL);
        (void)hipEventRecord(cx->t1, s);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpyAsync(outc, d_outc, sizeof(double) * P * KM, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(outsse, d_sse, sizeof(double) * P, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipMemcpyAsync(outstatus, d_st, sizeof(int) * P, hipMemcpyDeviceToHost, s) == hipSuccess &&
             hipStreamSynchronize(s) == hipSuccess;
    }
    float ms = -1.0f;
    if (ok) (void)hipEventElapsedTime(&ms, cx->t0, cx->t1);
    (void)hipGetLastError();
    release();
    return ok ? ms : -1.0f;
}

} /* extern "C" */
