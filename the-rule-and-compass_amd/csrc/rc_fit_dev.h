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
    int pad;
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

/* ---- f64 row evaluation of the post-order records ------------------------ */

/* Values of every node for one row in f64 (ft_forward_d, eval.cu:1324-1360):
   v[i] is node i's value, v[m - 1] the prediction. */
__device__ __forceinline__ void ft_forward64(const uint32_t *__restrict__ nd, int m,
                                             const double *__restrict__ cols64, int n, int row,
                                             const double *c, double *v)
{
    for (int i = 0; i < m; ++i) {
        const uint32_t op = nd[4 * i], k0 = nd[4 * i + 1], k1 = nd[4 * i + 2], slot = nd[4 * i + 3];
        if (op == OP_CONST) v[i] = c[slot];
        else if (op == OP_VAR) v[i] = cols64[(size_t)slot * (size_t)n + (size_t)row];
        else if (ft_is_binary(op)) v[i] = ft_binary(op, v[k0], v[k1]);
        else v[i] = ft_unary(op, v[k0]);
    }
}

/* Adjoint sweep of one row in f64 (ft_backward_d, eval.cu:1366-1463): g[j]
   = d prediction / d constant j for the p constants; a node whose adjoint is
   0 is skipped. */
__device__ __forceinline__ void ft_backward64(const uint32_t *__restrict__ nd, int m, const double *v,
                                              double *adj, double *g, int p)
{
    for (int i = 0; i < m; ++i) adj[i] = 0.0;
    for (int j = 0; j < p; ++j) g[j] = 0.0;
    adj[m - 1] = 1.0;
    for (int i = m - 1; i >= 0; --i) {
        const double a = adj[i];
        if (a == 0.0) continue;
        const uint32_t op = nd[4 * i], k0 = nd[4 * i + 1], k1 = nd[4 * i + 2], slot = nd[4 * i + 3];
        if (op == OP_CONST) { g[slot] += a; continue; }
        if (op == OP_VAR) continue;
        const double x = v[k0];
        if (op == OP_ADD) { adj[k0] += a; adj[k1] += a; }
        else if (op == OP_SUB) { adj[k0] += a; adj[k1] -= a; }
        else if (op == OP_MUL) { adj[k0] += a * v[k1]; adj[k1] += a * x; }
        else if (op == OP_DIV) { const double y = v[k1]; adj[k0] += a / y; adj[k1] += -a * x / (y * y); }
        else if (op == OP_POW) {
            const double y = v[k1];
            if (x > 0.0) { adj[k0] += a * y * pow(x, y - 1.0); adj[k1] += a * pow(x, y) * log(x); }
        } else {
            adj[k0] += a * ft_unary_partial<double>(op, x, v[i]);
        }
    }
}

/* Damped, Jacobi-equilibrated Cholesky solve of (A~ + lambda diag + 1e-12)
   delta~ = -D^-1/2 J^T r with the scale-free ln det gate on the undamped
   system (ft_solve, eval.cu:1874-1949).  jtj is the packed lower triangle
   (i*(i+1)/2 + j).  Returns 0 ok, 4 rank-deficient but solved, 1 no finite
   step.  Serial: run by one lane.  Every multiply-subtract is an explicit
   fma, so the wave-parallel form below produces the same bits. */
__device__ int ft_solve(const double *jtj, const double *jtr, double lambda, int p, double *delta,
                        double *L)
{
    double dscale[FT_KMAX];
    for (int i = 0; i < p; ++i) {
        const double aii = jtj[i * (i + 1) / 2 + i];
        if (!(aii > 0.0) || !isfinite(aii)) return 1;
        dscale[i] = sqrt(aii);
    }
    int rank_deficient = 0;
    double sum_ln_L = 0.0;
    int ok = 1;
    for (int i = 0; i < p && ok; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = jtj[i * (i + 1) / 2 + j] / (dscale[i] * dscale[j]);
            if (!isfinite(s)) return 1;
            for (int q = 0; q < j; ++q) s = fma(-L[i * FT_KMAX + q], L[j * FT_KMAX + q], s);
            if (i == j) {
                if (!(s > 0.0) || !isfinite(s)) { ok = 0; break; }
                L[i * FT_KMAX + i] = sqrt(s);
                sum_ln_L = fma(0.5, log(s), sum_ln_L);
            } else {
                L[i * FT_KMAX + j] = s / L[j * FT_KMAX + j];
            }
        }
    }
    if (!ok) {
        rank_deficient = 1;
    } else {
        const double ln_det = 2.0 * sum_ln_L;
        if (!isfinite(ln_det) || ln_det < FT_LN_DET_FLOOR) rank_deficient = 1;
    }
    for (int i = 0; i < p; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = jtj[i * (i + 1) / 2 + j] / (dscale[i] * dscale[j]);
            if (i == j) s = fma(s, 1.0 + lambda, 1e-12);
            for (int q = 0; q < j; ++q) s = fma(-L[i * FT_KMAX + q], L[j * FT_KMAX + q], s);
            if (i == j) {
                if (!(s > 1e-18) || !isfinite(s)) return 1;
                L[i * FT_KMAX + i] = sqrt(s);
            } else {
                L[i * FT_KMAX + j] = s / L[j * FT_KMAX + j];
            }
        }
    }
    double rhs[FT_KMAX];
    for (int i = 0; i < p; ++i) rhs[i] = -jtr[i] / dscale[i];
    for (int i = 0; i < p; ++i) {
        double s = rhs[i];
        for (int q = 0; q < i; ++q) s = fma(-L[i * FT_KMAX + q], rhs[q], s);
        rhs[i] = s / L[i * FT_KMAX + i];
    }
    for (int i = p - 1; i >= 0; --i) {
        double s = rhs[i];
        for (int q = i + 1; q < p; ++q) s = fma(-L[q * FT_KMAX + i], delta[q], s);
        delta[i] = s / L[i * FT_KMAX + i];
    }
    for (int i = 0; i < p; ++i) delta[i] /= dscale[i];
    return rank_deficient ? 4 : 0;
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

/* One Cholesky factorization of the scaled lower triangle in S (row-major,
   stride FT_KMAX), right-looking over the 64 lanes of a one-wave workgroup:
   at step j the diagonal is checked and square-rooted, column j is divided
   by it and the trailing triangle receives its q = j update.  Each entry thus
   gets ft_solve's updates in increasing q, then its divide or root: the same
   bits as the serial row-by-row order.  Returns the first row whose diagonal
   fails (p when none); lane 0 accumulates 0.5 ln s of the accepted diagonals
   in row order. */
template <bool UNDAMPED>
__device__ int ft_chol_wave(double *S, int p, int lane, double &sum_ln)
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
        for (int t = lane; t < n * (n + 1) / 2; t += 64) {
            int a, b;
            ft_tri_index(t, a, b);
            const int i = j + 1 + a, m = j + 1 + b;
            S[i * FT_KMAX + m] = fma(-S[i * FT_KMAX + j], S[m * FT_KMAX + j], S[i * FT_KMAX + m]);
        }
        __syncthreads();
    }
    return p;
}

/* ft_solve run by the 64 lanes of a one-wave workgroup, with the same result
   bits and status: both factorizations are wave-parallel (ft_chol_wave), the
   serial solve's early exits are resolved in its order (a non-finite scaled
   entry in a row the serial pass reaches returns 1 before that row's
   diagonal is judged), and the two triangular solves run on lane 0 in
   ft_solve's order.  jtj, jtr in any memory; S (FT_KMAX^2), dsc and rhs
   (FT_KMAX) are LDS scratch; delta (LDS, FT_KMAX) is complete in every lane
   on return.  The status is the same in every lane. */
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
    unsigned nfrow = 0; /* rows holding a non-finite scaled entry */
    for (int t = lane; t < ntri; t += 64) {
        int i, j;
        ft_tri_index(t, i, j);
        const double s = jtj[t] / (dsc[i] * dsc[j]);
        if (!isfinite(s)) nfrow |= 1u << i;
        S[i * FT_KMAX + j] = s;
    }
    for (int off = 32; off > 0; off >>= 1) nfrow |= (unsigned)__shfl_xor((int)nfrow, off, 64);
    __syncthreads();
    double sum_ln = 0.0;
    const int fail = ft_chol_wave<true>(S, p, lane, sum_ln);
    const unsigned reached = fail >= p ? 0xffffffffu : (fail >= 31 ? 0xffffffffu : ((2u << fail) - 1u));
    if (nfrow & reached) return 1;
    int rank_deficient = 1;
    if (fail >= p) {
        const double ln_det = 2.0 * __shfl(sum_ln, 0, 64);
        rank_deficient = (!isfinite(ln_det) || ln_det < FT_LN_DET_FLOOR) ? 1 : 0;
    }
    __syncthreads();
    for (int t = lane; t < ntri; t += 64) {
        int i, j;
        ft_tri_index(t, i, j);
        double s = jtj[t] / (dsc[i] * dsc[j]);
        if (i == j) s = fma(s, 1.0 + lambda, 1e-12);
        S[i * FT_KMAX + j] = s;
    }
    __syncthreads();
    double unused = 0.0;
    if (ft_chol_wave<false>(S, p, lane, unused) < p) return 1;
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
