/*
 * Device pieces of the batched LM / IRLS fit (included by rc_fit.hip).
 *
 * Program form.  The ABI's post-order records [op, kid0, kid1, slot]
 * (fit.rs:1101-1196) are translated on the host into two wave-uniform
 * instruction streams per candidate:
 *   forward  one word per node: op | flags << 8 | arg << 16, executed as a
 *            stack machine; flags ask the node to push the local partial
 *            derivatives its on-path children need onto the tape;
 *   reverse  one word per node on a path to a constant (or an input, for the
 *            input gradient), in reverse post-order: the adjoint sweep of
 *            ft_backward (eval.cu:1236-1317) becomes an adjoint *stack*
 *            (a tree node has one parent, so the next on-path node in
 *            reverse post-order always consumes the top of the stack) and the
 *            tape of partials is a LIFO in the same order.
 * Value stack, adjoint stack and tape live in VGPRs as LLVM vector values
 * indexed by wave-uniform positions (s_set_gpr_idx), so neither sweep touches
 * scratch or LDS.
 */
#ifndef RC_FIT_DEV_H
#define RC_FIT_DEV_H

#include "rc_device.h"

#define FT_KMAX 20
#define FT_MAX_NODES 96
#define FT_TRI_MAX 253 /* (20 + 2)(20 + 3) / 2: [J | r_hi | r_lo] lower triangle */

/* Forward flags. */
#define FF_D0 1u /* push the partial w.r.t. the left (or only) child */
#define FF_D1 2u /* push the partial w.r.t. the right child */

/* Reverse kinds. */
#define RK_CONST 0u
#define RK_VAR 1u
#define RK_ADD 2u
#define RK_SUB 3u
#define RK_NEG 4u
#define RK_UNARY 5u
#define RK_BINARY 6u
#define RF_LEFT 1u
#define RF_RIGHT 2u

/* Per-candidate descriptor (host-built). */
struct FtCand {
    uint32_t fwd_off;
    uint32_t rev_off;
    int m;       /* forward words */
    int nrev;    /* reverse words */
    int k;       /* constants */
    int rowset;
    int metric;
    int item0;   /* first work item */
    int nitems;
    int vm;      /* 1: fit VM code (word pairs, m / nrev count pairs) */
    uint32_t t_off; /* value-pass stream (trial and exact losses) */
    int t_m;        /* its instruction count */
};

/* Gram entries of a column pairing for K <= S + 16 <= 24 columns in two
   16 x 16 products per four rows (where a blockwise 32 x 32 Gram takes
   three): P pairs columns 0..15 (rows) with columns S..S+15 (columns), Q
   pairs the columns c(i) = i < 8 ? i : 8 + i with themselves (S <= 8).
   Triangle entry (a, b), a >= b, comes from P (row b, column a - S) when
   a >= 8 and b < 16, else from Q (both below 8 or both at least 16), so
   every entry is stored once and is the same dot product over the same rows
   as in the blockwise Gram.  (i, j) is the caller's position in both
   products. */
template <int S>
__device__ __forceinline__ void ft_pair_store(double *pg, int i, int j, int K, double p, double q)
{
    static_assert(S >= 0 && S <= 8, "Q covers the columns below 8 and from 16");
    const int a = S + j;
    if (a >= 8 && i <= a && a < K) pg[a * (a + 1) / 2 + i] = p;
    const int ci = i < 8 ? i : 8 + i, cj = j < 8 ? j : 8 + j;
    if (i >= j && ci < K && (ci < 8 || cj >= 16)) pg[ci * (ci + 1) / 2 + cj] = q;
}

/* One resident row set as the kernels see it. */
struct FtRows {
    const float *cols;
    const double *cols64;
    const double *y64;
    const double *w64;
    int n;
    int v;
};

/* One work item: a tile of one candidate's rows. */
struct FtItem {
    int cand;
    int row0;
    int rows;
    int p64_blk; /* start of the item's rows in the f64 prediction buffer, in 64-row blocks */
};

typedef float fv8 __attribute__((ext_vector_type(8)));
typedef float fv16 __attribute__((ext_vector_type(16)));
typedef float fv32 __attribute__((ext_vector_type(32)));
typedef double dv8 __attribute__((ext_vector_type(8)));
typedef double dv16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

/* ---- value semantics ----------------------------------------------------- */

/* f32: the reference device forward (ft_forward, eval.cu:1173-1233). */
__device__ __forceinline__ float ft_unary(unsigned op, float a)
{
    switch (op) {
        case OP_SIN: return sinf(a);
        case OP_COS: return cosf(a);
        case OP_TAN: { float t = tanf(a); return isfinite(t) ? t : RC_F32_NAN; }
        case OP_EXP: return (a > 700.0f) ? __builtin_inff() : expf(a);
        case OP_LOG: return (a > 0.0f) ? logf(a) : RC_F32_NAN;
        case OP_SQRT: return (a >= 0.0f) ? sqrtf(a) : RC_F32_NAN;
        case OP_ABS: return fabsf(a);
        case OP_TANH: return tanhf(a);
        case OP_SIG: return rc_logistic_f(a);
        case OP_SQUARE: return a * a;
        case OP_CUBE: return a * a * a;
        case OP_NEG: return -a;
        case OP_GAUSS: return expf(-a * a);
        case OP_SINH: return sinhf(a);
        case OP_COSH: return coshf(a);
        case OP_ERF: return (a == 0.0f) ? 0.0f : erff(a);
        case OP_ERFC: return erfcf(a);
        default: return RC_F32_NAN;
    }
}

__device__ __forceinline__ float ft_binary(unsigned op, float a, float b)
{
    switch (op) {
        case OP_ADD: return a + b;
        case OP_SUB: return a - b;
        case OP_MUL: return a * b;
        case OP_DIV: return (fabsf(b) < 1e-12f) ? RC_F32_NAN : a / b;
        case OP_POW: return (a < 0.0f && fabsf(b - roundf(b)) > 1e-9f) ? RC_F32_NAN : powf(a, b);
        default: return RC_F32_NAN;
    }
}

/* f64: the CPU fitter's value_of (fit.rs:1211-1365). */
__device__ __forceinline__ double ft_unary(unsigned op, double a)
{
    switch (op) {
        case OP_SIN: return sin(a);
        case OP_COS: return cos(a);
        case OP_TAN: return tan(a);
        case OP_EXP: return (a > 700.0) ? __builtin_inf() : exp(a);
        case OP_LOG: return log(a);
        case OP_SQRT: return sqrt(a);
        case OP_ABS: return fabs(a);
        case OP_TANH: return tanh(a);
        case OP_SIG: return 1.0 / (1.0 + exp(-a));
        case OP_SQUARE: return a * a;
        case OP_CUBE: return a * a * a;
        case OP_NEG: return -a;
        case OP_GAUSS: return exp(-a * a);
        case OP_SINH: return sinh(a);
        case OP_COSH: return cosh(a);
        case OP_ERF: return erf(a);
        case OP_ERFC: return erfc(a);
        default: return RC_F64_NAN;
    }
}

__device__ __forceinline__ double ft_binary(unsigned op, double a, double b)
{
    switch (op) {
        case OP_ADD: return a + b;
        case OP_SUB: return a - b;
        case OP_MUL: return a * b;
        case OP_DIV: return (fabs(b) < 1e-12) ? RC_F64_NAN : a / b;
        case OP_POW: return pow(a, b);
        default: return RC_F64_NAN;
    }
}

__device__ __forceinline__ bool ft_is_binary(unsigned op)
{
    return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV || op == OP_POW;
}

/* Local partial of a unary operator at operand x with value `out`; the same
   derivative table as ft_backward (eval.cu:1261-1315). */
template <typename T>
__device__ __forceinline__ T ft_unary_partial(unsigned op, T x, T out)
{
    switch (op) {
        case OP_NEG: return (T)-1;
        case OP_EXP: return out;
        case OP_LOG: return (T)1 / x;
        case OP_ABS: return x >= (T)0 ? (T)1 : (T)-1;
        case OP_SQRT: return (T)1 / ((T)2 * out);
        case OP_SIN: return cos(x);
        case OP_COS: return -sin(x);
        case OP_TAN: { T c = cos(x); return (T)1 / (c * c); }
        case OP_SINH: return cosh(x);
        case OP_COSH: return sinh(x);
        case OP_TANH: return (T)1 - out * out;
        case OP_SIG: return out * ((T)1 - out);
        case OP_GAUSS: return (T)-2 * x * out;
        case OP_SQUARE: return (T)2 * x;
        case OP_CUBE: return (T)3 * x * x;
        case OP_ERF: return (T)1.1283791670955126 * exp(-x * x);
        case OP_ERFC: return (T)-1.1283791670955126 * exp(-x * x);
        default: return (T)0;
    }
}

/* ---- the solve ----------------------------------------------------------- */

/* ln(1e-12): the smallest identified fraction of parameter space. */
#define FT_LN_DET_FLOOR (-27.631021115928547)

/* ---- f64 sweeps over the post-order records, one row per thread ---------- */

/* One record [op, left, right, slot] of a post-order fit program. */
struct FtRec {
    uint32_t op, left, right, slot;
};
__device__ __forceinline__ FtRec ft_rec(const uint32_t *__restrict__ recs, int i)
{
    const uint32_t *w = recs + 4 * i;
    return {w[0], w[1], w[2], w[3]};
}

/* value_of (fit.rs:1211-1365) for every record of one row: val[i] is record
   i's value, val[m - 1] the prediction. */
__device__ __forceinline__ void ft_rec_values(const uint32_t *__restrict__ recs, int m,
                                              const double *__restrict__ cols64, int n, int row,
                                              const double *c, double *val)
{
    for (int i = 0; i < m; ++i) {
        const FtRec r = ft_rec(recs, i);
        double out;
        switch (r.op) {
            case OP_CONST: out = c[r.slot]; break;
            case OP_VAR: out = cols64[(size_t)r.slot * (size_t)n + (size_t)row]; break;
            default:
                out = ft_is_binary(r.op) ? ft_binary(r.op, val[r.left], val[r.right])
                                         : ft_unary(r.op, val[r.left]);
        }
        val[i] = out;
    }
}

/* What a record with adjoint `a` hands to its children (fit.rs:1503-1569):
   to_l for the left (or only) child, to_r for the right one.  Each
   expression keeps the CPU fitter's operation order, so +, -, * and /
   programs give the oracle's bits. */
__device__ __forceinline__ void ft_rec_hand_down(uint32_t op, double a, double l, double r,
                                                 double self, double &to_l, double &to_r)
{
    to_r = 0.0;
    switch (op) {
        case OP_ADD: to_l = a; to_r = a; break;
        case OP_SUB: to_l = a; to_r = -a; break;
        case OP_MUL: to_l = a * r; to_r = a * l; break;
        case OP_DIV: to_l = a / r; to_r = -a * l / (r * r); break;
        case OP_POW:
            to_l = 0.0;
            if (l > 0.0) {
                to_l = a * r * pow(l, r - 1.0);
                to_r = a * pow(l, r) * log(l);
            }
            break;
        default: to_l = a * ft_unary_partial<double>(op, l, self); break;
    }
}

/* Reverse accumulation over the records of one row after ft_rec_values:
   grad[j] = d prediction / d constant j for the p constants.  adj is scratch
   for m values; a record whose adjoint is 0 hands nothing down. */
__device__ __forceinline__ void ft_rec_gradient(const uint32_t *__restrict__ recs, int m,
                                                const double *val, double *adj, double *grad,
                                                int p)
{
    for (int j = 0; j < p; ++j) grad[j] = 0.0;
    for (int i = 0; i + 1 < m; ++i) adj[i] = 0.0;
    adj[m - 1] = 1.0;
    for (int i = m - 1; i >= 0; --i) {
        const double a = adj[i];
        if (a == 0.0) continue;
        const FtRec r = ft_rec(recs, i);
        if (r.op == OP_CONST) {
            grad[r.slot] += a;
        } else if (r.op != OP_VAR) {
            double to_l, to_r;
            const bool two = ft_is_binary(r.op);
            ft_rec_hand_down(r.op, a, val[r.left], two ? val[r.right] : 0.0, val[i], to_l, to_r);
            adj[r.left] += to_l;
            if (two) adj[r.right] += to_r;
        }
    }
}

/* (i, j), j <= i, of packed lower-triangle index t = i (i + 1) / 2 + j. */
__device__ __forceinline__ void ft_tri_index(int t, int &i, int &j)
{
    int r = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    i = r;
    j = t - r * (r + 1) / 2;
}

/* (i, j) of the packed lower-triangle indices t = lane + 64 u, u < 4: every
   entry of a triangle of up to FT_KMAX rows, worked out once per solve. */
struct FtTriMap {
    int i[4], j[4];
};
static_assert(FT_KMAX * (FT_KMAX + 1) / 2 <= 4 * 64, "four triangle entries per lane");
__device__ __forceinline__ FtTriMap ft_tri_map(int lane)
{
    FtTriMap m;
#pragma unroll
    for (int u = 0; u < 4; ++u) ft_tri_index(lane + 64 * u, m.i[u], m.j[u]);
    return m;
}

/* One Cholesky factorization of the scaled lower triangle in S (row-major,
   stride FT_KMAX), right-looking over the 64 lanes of a one-wave workgroup:
   at step j the diagonal is checked and square-rooted, column j is divided
   by it and the trailing triangle receives its q = j update.  Each entry
   thus receives its fma updates in increasing q and then its divide or root,
   so the factor does not depend on how the lanes share the work.  Returns
   the first row whose diagonal fails (p when none); lane 0 accumulates
   0.5 ln s of the accepted diagonals in row order. */
template <bool UNDAMPED>
__device__ int ft_chol_wave(double *S, int p, int lane, double &sum_ln, const FtTriMap &tm)
{
    for (int j = 0; j < p; ++j) {
        const double d = S[j * FT_KMAX + j];
        const bool good = UNDAMPED ? (d > 0.0 && isfinite(d)) : (d > 1e-18 && isfinite(d));
        if (!good) return j; /* d is the same LDS word in every lane */
        const double ljj = sqrt(d);
        __syncthreads();
        if (lane == 0) {
            S[j * FT_KMAX + j] = ljj;
            if (UNDAMPED) sum_ln = fma(0.5, log(d), sum_ln);
        }
        for (int i = j + 1 + lane; i < p; i += 64) S[i * FT_KMAX + j] = S[i * FT_KMAX + j] / ljj;
        __syncthreads();
        const int n = p - j - 1;
        const int nt = n * (n + 1) / 2;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (lane + 64 * u >= nt) break;
            const int i = j + 1 + tm.i[u], m = j + 1 + tm.j[u];
            S[i * FT_KMAX + m] = fma(-S[i * FT_KMAX + j], S[m * FT_KMAX + j], S[i * FT_KMAX + m]);
        }
        __syncthreads();
    }
    return p;
}

/* The LM step solve (eval.cu:1874-1949) on the 64 lanes of a one-wave
   workgroup: Jacobi equilibration by the diagonal of the packed lower
   triangle jtj (i (i + 1) / 2 + j), the scale-free ln det gate on the
   undamped equilibrated system (rank deficient below ln 1e-12), then the
   damped system s (1 + lambda) + 1e-12 on the diagonal factorized and solved
   for delta = -D^-1/2 (A~ + damping)^-1 D^-1/2 J^T r.  Returns 0, 4 (rank
   deficient but solved) or 1 (no finite step: a non-positive or non-finite
   diagonal, a non-finite scaled entry in a row the factorization reaches
   before failing, or a damped pivot at or below 1e-18).  Both
   factorizations are wave-parallel (ft_chol_wave); the triangular solves run
   on lane 0 in row order with explicit fma.  jtj, jtr in any memory; S
   (FT_KMAX^2), dsc and rhs (FT_KMAX) are LDS scratch; delta (LDS, FT_KMAX)
   is complete in every lane on return.  The status is the same in every
   lane. */
__device__ int ft_solve_wave(const double *jtj, const double *jtr, double lambda, int p,
                             double *delta, double *S, double *dsc, double *rhs, int lane)
{
    int bad = 0;
    if (lane < p) {
        const double aii = jtj[lane * (lane + 1) / 2 + lane];
        if (!(aii > 0.0) || !isfinite(aii)) bad = 1;
        else dsc[lane] = sqrt(aii);
    }
    if (__any(bad)) return 1;
    __syncthreads();
    const int ntri = p * (p + 1) / 2;
    const FtTriMap tm = ft_tri_map(lane);
    unsigned nfrow = 0; /* rows holding a non-finite scaled entry */
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = lane + 64 * u;
        if (t >= ntri) break;
        const int i = tm.i[u], j = tm.j[u];
        const double s = jtj[t] / (dsc[i] * dsc[j]);
        if (!isfinite(s)) nfrow |= 1u << i;
        S[i * FT_KMAX + j] = s;
    }
    for (int off = 32; off > 0; off >>= 1) nfrow |= (unsigned)__shfl_xor((int)nfrow, off, 64);
    __syncthreads();
    double sum_ln = 0.0;
    const int fail = ft_chol_wave<true>(S, p, lane, sum_ln, tm);
    const unsigned reached = fail >= p ? 0xffffffffu : (fail >= 31 ? 0xffffffffu : ((2u << fail) - 1u));
    if (nfrow & reached) return 1;
    int rank_deficient = 1;
    if (fail >= p) {
        const double ln_det = 2.0 * __shfl(sum_ln, 0, 64);
        rank_deficient = (!isfinite(ln_det) || ln_det < FT_LN_DET_FLOOR) ? 1 : 0;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int t = lane + 64 * u;
        if (t >= ntri) break;
        const int i = tm.i[u], j = tm.j[u];
        double s = jtj[t] / (dsc[i] * dsc[j]);
        if (i == j) s = fma(s, 1.0 + lambda, 1e-12);
        S[i * FT_KMAX + j] = s;
    }
    __syncthreads();
    double unused = 0.0;
    if (ft_chol_wave<false>(S, p, lane, unused, tm) < p) return 1;
    if (lane == 0) {
        for (int i = 0; i < p; ++i) {
            double s = -jtr[i] / dsc[i];
            for (int q = 0; q < i; ++q) s = fma(-S[i * FT_KMAX + q], rhs[q], s);
            rhs[i] = s / S[i * FT_KMAX + i];
        }
        for (int i = p - 1; i >= 0; --i) {
            double s = rhs[i];
            for (int q = i + 1; q < p; ++q) s = fma(-S[q * FT_KMAX + i], delta[q], s);
            delta[i] = s / S[i * FT_KMAX + i];
        }
        for (int i = 0; i < p; ++i) delta[i] /= dsc[i];
    }
    __syncthreads();
    return rank_deficient ? 4 : 0;
}

#endif
