/*
 * Batched Levenberg-Marquardt / IRLS constant fitting on gfx950.
 *
 * Reference behaviour: fit_kernel (eval.cu:2265-2606) with ft_forward /
 * ft_backward (1173-1317), ft_assemble (1693-1804), ft_loss (1809-1855),
 * ft_solve (1874-1949), the f64 re-baseline and Gauss-Newton polish
 * (2536-2580), the status labels (2583-2605) and rc_fit_batch (4001-4078).
 *
 * MI355X design.  The reference runs one thread block per candidate for the
 * whole fit, so a candidate's rows are limited to one SM and uneven candidates
 * leave SMs idle.  Here every LM iteration is a short sequence of batch-wide
 * launches over work items (candidate, 8192-row tile), so all 256 CUs share
 * every pass whatever the mix of program lengths, constant counts and row
 * counts:
 *   ft_asm_kernel    f32 forward + adjoint per row (one wave per item), the
 *                    augmented row [s*J | s*r_hi | s*r_lo] written to LDS and
 *                    reduced on FP32 MFMA (v_mfma_f32_16x16x4f32): J^T J,
 *                    J^T r and r^T r in one product.  r = pred - y is formed in
 *                    f64 and split hi/lo so J^T r keeps f64-level accuracy;
 *                    the MFMA accumulators are promoted to f64 every 256 rows;
 *   ft_trial_kernel  forward only: sum w*rho(r) in f64 (f32 rows in the LM
 *                    loop, f64 rows for the exact re-baseline and the polish);
 *   ft_ctrl_kernel   one wave per candidate: fixed-order reduction of the
 *                    item partials, the FP64 damped solve, acceptance, lambda
 *                    schedule, IRLS rounds and status labels.
 * Partials are per (candidate, tile) and reduced in tile order, so a
 * candidate's result depends only on its own program and rows, never on the
 * batch it came in.  The launch sequence is recorded once per batch shape as
 * a hipGraph and replayed.
 */
#include "rc_fit_dev.h"
#include "rc_fitvm.h"
#include "rc_fitvm_kern.h"
#include "rc_internal.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int XS = 66;          /* LDS column stride of the augmented rows */
constexpr int TILE_ROWS = 8192; /* built-in rows per work item (a multiple of 256) */
/* Largest f64 prediction buffer of a fit (bytes): 8 bytes per row of every
   candidate; 1024 candidates over 262 144 rows take 2 GiB. */
constexpr size_t FT_P64_CAP = (size_t)32 << 30;
/* Interpreter capacities.  Every vector update is unconditional (the op's
   arithmetic runs on scalars), so a slot above the live area is written each
   step: the small class needs depth + 1 value slots, adjoint depth + 2 slots
   and 31 entries per tape lane. */
constexpr int SMALL_S = 8, SMALL_AS = 8, SMALL_TAPE = 32;
constexpr int BIG_S = 100, BIG_AS = 100, BIG_TAPE = 100;

/* Candidate stages of the device state machine. */
enum {
    ST_NEED_ASM = 1, ST_NEED_TRIAL = 2, ST_LM_DONE = 3, ST_P_NEED_ASM = 4, ST_P_NEED_TRIAL = 5,
    ST_DONE = 6
};

/* Control operations. */
enum {
    CT_INIT = 0, CT_AFTER_ASM = 1, CT_AFTER_TRIAL = 2, CT_PRE_REBASE = 3, CT_REBASE = 4,
    CT_AFTER_ASM_P = 5, CT_AFTER_TRIAL_P = 6, CT_FINAL = 7
};

struct FtState {
    double c[FT_KMAX];
    double trial[FT_KMAX];
    double jtj[FT_KMAX * (FT_KMAX + 1) / 2];
    double jtr[FT_KMAX];
    double best;
    double lambda;
    int k, metric, irls, rounds, inner, iters;
    int stage, it, round, assembled, first;
    int status, sstart, srank, improved, skipped;
    int halvings, polish_it, pre_status, acc_bits; /* acc_bits: LM decision q accepted (bit q) */
    int n_asm, n_trial, n_exact, n_decided; /* passes consumed, LM decisions (telemetry) */
    int pruned; /* this LM trial was rejected on its first half (ft_trial_prune) */
    int pad_[2];
};
static_assert(sizeof(FtState) % 16 == 0, "FtState is copied in 16-byte words");

/* Wave-uniform storage for the interpreter: an LLVM vector (registers,
   indexed through s_set_gpr_idx) for the common sizes, a plain array for the
   rare oversized program.  Callers update them unconditionally (one indexed
   read and one indexed write per step) so no branch join ever merges two
   versions of a vector. */
template <typename T, int N> struct VStore;
template <> struct VStore<float, 8> { fv8 v; __device__ float get(int i) const { return v[i]; } __device__ void set(int i, float x) { v[i] = x; } };
template <> struct VStore<float, 32> { fv32 v; __device__ float get(int i) const { return v[i]; } __device__ void set(int i, float x) { v[i] = x; } };
template <> struct VStore<double, 8> { dv8 v; __device__ double get(int i) const { return v[i]; } __device__ void set(int i, double x) { v[i] = x; } };
typedef double dv32 __attribute__((ext_vector_type(32)));
template <> struct VStore<double, 32> { dv32 v; __device__ double get(int i) const { return v[i]; } __device__ void set(int i, double x) { v[i] = x; } };
template <typename T, int N> struct AStore {
    T a[N];
    __device__ T get(int i) const { return a[i]; }
    __device__ void set(int i, T x) { a[i] = x; }
};

template <typename T, bool BIG> struct Stores;
template <> struct Stores<float, false> {
    typedef VStore<float, SMALL_S> Val;
    typedef VStore<float, SMALL_AS> Adj;
    typedef VStore<float, SMALL_TAPE> Tape;
};
template <> struct Stores<float, true> {
    typedef AStore<float, BIG_S> Val;
    typedef AStore<float, BIG_AS> Adj;
    typedef AStore<float, BIG_TAPE> Tape;
};
/* The f64 tapes serve the plain FP64 arm's assembly; the f64 value passes
   never touch them. */
template <> struct Stores<double, false> {
    typedef VStore<double, SMALL_S> Val;
    typedef VStore<double, SMALL_AS> Adj;
    typedef VStore<double, SMALL_TAPE> Tape;
};
template <> struct Stores<double, true> {
    typedef AStore<double, BIG_S> Val;
    typedef AStore<double, BIG_AS> Adj;
    typedef AStore<double, BIG_TAPE> Tape;
};

template <typename T> struct XVec;
template <> struct XVec<float> { typedef fv8 type; };
template <> struct XVec<double> { typedef dv8 type; };

/* The two LIFO lanes of the tape: partials w.r.t. a left (or only) child and
   w.r.t. a right child.  Each lane's pushes and pops mirror each other. */
template <typename T, bool BIG> struct Tape2 {
    typename Stores<T, BIG>::Tape l, r;
    int tl = 0, tr = 0;
};

/* Forward sweep of one row (stack machine, top of stack cached).  With TAPE,
   on-path nodes push their local partials (FF_D0 / FF_D1). */
template <typename T, bool BIG, bool TAPE>
__device__ __forceinline__ T ft_forward(const uint32_t *__restrict__ fc, int m,
                                        const T *__restrict__ cst,
                                        const typename XVec<T>::type &xv,
                                        const T *__restrict__ cols, int n, int row,
                                        Tape2<T, BIG> &tape)
{
    typename Stores<T, BIG>::Val st;
    T tos = (T)0;
    int sp = 0; /* st[1 .. sp-1] hold the values below the top; st[0] is scratch */
    for (int i = 0; i < m; ++i) {
        const uint32_t w = fc[i];
        const uint32_t op = w & 0xffu;
        const uint32_t fl = (w >> 8) & 0xffu;
        const uint32_t arg = w >> 16;
        const T a = st.get(sp > 0 ? sp - 1 : 0);
        const T b = tos;
        T res, d0 = (T)0, d1 = (T)0;
        int push = 0, pop = 0;
        if (op == OP_CONST) {
            res = cst[arg];
            push = 1;
        } else if (op == OP_VAR) {
            res = (arg < 8u) ? xv[arg] : cols[(size_t)arg * (size_t)n + (size_t)row];
            push = 1;
        } else if (op == OP_MUL) {
            res = a * b;
            d0 = b;
            d1 = a;
            pop = 1;
        } else if (op == OP_ADD) {
            res = a + b;
            pop = 1;
        } else if (op == OP_SUB) {
            res = a - b;
            pop = 1;
        } else if (op == OP_DIV || op == OP_POW) {
            res = ft_binary(op, a, b);
            if (TAPE && fl) {
                if (op == OP_DIV) { d0 = (T)1 / b; d1 = -a / (b * b); }
                else if (a > (T)0) { d0 = b * pow(a, b - (T)1); d1 = pow(a, b) * log(a); }
            }
            pop = 1;
        } else {
            res = ft_unary(op, b);
            if (TAPE && fl) d0 = ft_unary_partial<T>(op, b, res);
        }
        st.set(sp, b);          /* the old top: kept only on a push */
        sp += push - pop;
        tos = res;
        if (TAPE) {
            tape.l.set(tape.tl, d0);
            tape.r.set(tape.tr, d1);
            tape.tl += (fl & FF_D0) ? 1 : 0;
            tape.tr += (fl & FF_D1) ? 1 : 0;
        }
    }
    return tos;
}

/* Reverse sweep seeded with `seed` at the root.  Constant gradients go to
   the LDS column block X (X[slot * XSTR + lane], in the block's element
   type); input gradients, when asked, to gx.  Returns whether every constant
   gradient is finite. */
template <typename T, bool BIG, typename XT = float, int XSTR = XS>
__device__ __forceinline__ bool ft_reverse(const uint32_t *__restrict__ rc, int nrev, T seed,
                                           Tape2<T, BIG> &tape, XT *X, int lane, T *gx)
{
    typename Stores<T, BIG>::Adj ad;
    ad.set(0, seed);
    int ap = 1;
    bool fin = true;
    for (int j = 0; j < nrev; ++j) {
        const uint32_t w = rc[j];
        const uint32_t kind = w & 0xffu;
        const uint32_t fl = (w >> 8) & 0xffu;
        const uint32_t arg = w >> 16;
        const T a = ad.get(ap - 1);
        const T dl = tape.l.get(tape.tl > 0 ? tape.tl - 1 : 0);
        const T dr = tape.r.get(tape.tr > 0 ? tape.tr - 1 : 0);
        const bool z = (a == (T)0);
        T first = (T)0, second = (T)0;
        int npush = 0;
        if (kind == RK_CONST) {
            if (X) X[arg * XSTR + lane] = (XT)a;
            fin = fin && isfinite(a);
        } else if (kind == RK_VAR) {
            if (gx) gx[arg] += a;
        } else if (kind == RK_ADD || kind == RK_SUB) {
            const T ar = (kind == RK_SUB) ? -a : a;
            if ((fl & RF_LEFT) && (fl & RF_RIGHT)) { first = a; second = ar; npush = 2; }
            else if (fl & RF_LEFT) { first = a; npush = 1; }
            else { first = ar; npush = 1; }
        } else if (kind == RK_NEG) {
            first = -a;
            npush = 1;
        } else if (kind == RK_UNARY) {
            first = z ? (T)0 : a * dl;
            npush = 1;
        } else {
            const T gl = z ? (T)0 : a * dl;
            const T gr = z ? (T)0 : a * dr;
            if ((fl & RF_LEFT) && (fl & RF_RIGHT)) { first = gl; second = gr; npush = 2; }
            else if (fl & RF_LEFT) { first = gl; npush = 1; }
            else { first = gr; npush = 1; }
        }
        ad.set(ap - 1, first);
        ad.set(ap, second);
        ap += npush - 1;
        tape.tl -= (kind == RK_UNARY || (kind == RK_BINARY && (fl & RF_LEFT))) ? 1 : 0;
        tape.tr -= (kind == RK_BINARY && (fl & RF_RIGHT)) ? 1 : 0;
    }
    return fin;
}

__device__ __forceinline__ void load_x(const float *cols, int n, int v, int row, fv8 &xv)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) xv[q] = (q < v) ? cols[(size_t)q * (size_t)n + (size_t)row] : 0.0f;
}

__device__ __forceinline__ void load_x(const double *cols, int n, int v, int row, dv8 &xv)
{
#pragma unroll
    for (int q = 0; q < 8; ++q) xv[q] = (q < v) ? cols[(size_t)q * (size_t)n + (size_t)row] : 0.0;
}

template <typename T>
__device__ __noinline__ T ft_unary_call(unsigned op, T a) { return ft_unary(op, a); }
template <typename T>
__device__ __noinline__ T ft_binary_call(unsigned op, T a, T b) { return ft_binary(op, a, b); }

/* Value-only forward of R rows per lane (rows xrow[r]): the evaluator's
   interpreter shape.  The top of stack is cached, the values below it live
   in st[r][d] (VGPRs indexed by the wave-uniform depth while R * S is small),
   the common operators are inlined and the rare ones are out-of-line calls, so
   the scalar dispatch of each instruction is amortised over R rows. */
template <typename T, int S, int R>
__device__ __forceinline__ void ft_value_rows(const uint32_t *__restrict__ fc, int m,
                                              const T *__restrict__ cst,
                                              const T *__restrict__ cols, int n,
                                              const int *xrow, T *res)
{
    T tos[R];
    T st[R][S];
    int sp = 0, depth = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) tos[r] = (T)0;
    for (int i = 0; i < m; ++i) {
        const uint32_t w = fc[i];
        const uint32_t op = w & 0xffu;
        if (op == OP_VAR || op == OP_CONST) {
            if (depth > 0) {
#pragma unroll
                for (int r = 0; r < R; ++r) st[r][sp] = tos[r];
                ++sp;
            }
            ++depth;
            if (op == OP_VAR) {
                const T *col = cols + (size_t)(w >> 16) * (size_t)n;
#pragma unroll
                for (int r = 0; r < R; ++r) tos[r] = col[xrow[r]];
            } else {
                const T c = cst[w >> 16];
#pragma unroll
                for (int r = 0; r < R; ++r) tos[r] = c;
            }
        } else if (op == OP_MUL) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] * tos[r];
        } else if (op == OP_ADD) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] + tos[r];
        } else if (op == OP_SUB) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = st[r][sp] - tos[r];
        } else if (op == OP_DIV) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const T b = tos[r];
                tos[r] = (fabs(b) < (T)1e-12) ? (T)RC_F64_NAN : st[r][sp] / b;
            }
        } else if (op == OP_POW) {
            --sp; --depth;
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = ft_binary_call<T>(op, st[r][sp], tos[r]);
        } else if (op == OP_EXP) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = (tos[r] > (T)700) ? (T)__builtin_inf() : exp(tos[r]);
        } else if (op == OP_SIN) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = sin(tos[r]);
        } else if (op == OP_COS) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = cos(tos[r]);
        } else if (op == OP_NEG) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = -tos[r];
        } else if (op == OP_SQUARE) {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = tos[r] * tos[r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) tos[r] = ft_unary_call<T>(op, tos[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) res[r] = tos[r];
}

/* ---- assembly: forward + adjoint + MFMA Gram ------------------------------ */

/* POLISH: the exact polish's assembly.  The residual comes from an f64
   forward pass on the f64 rows (the Jacobian stays f32), so the iteration's
   fixed point is the f64 optimum; a non-finite gradient row is skipped and
   counted instead of failing the candidate (ft_assemble_d semantics). */
template <int KC, bool BIG, bool POLISH>
__global__ __launch_bounds__(64) void ft_asm_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run, const float *__restrict__ cp32,
    const double *__restrict__ cp64, double *__restrict__ pgram, int tri_stride,
    double *__restrict__ ploss, int *__restrict__ pbad, int *__restrict__ pskip)
{
    const bool polish = POLISH;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const FtCand d = cands[cand];
    const FtRows R = rowsd[d.rowset];
    const int lane = threadIdx.x;
    const int k = d.k;

    __shared__ float X[KC * XS];
    for (int i = lane; i < KC * XS; i += 64) X[i] = 0.0f;
    __syncthreads();

    const float *cst = cp32 + (size_t)cand * FT_KMAX;
    const uint32_t *fc = code + d.fwd_off;
    const uint32_t *rc = code + d.rev_off;
    floatx4 acc00 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
    double g00[4] = {0, 0, 0, 0}, g10[4] = {0, 0, 0, 0}, g11[4] = {0, 0, 0, 0};
    double lsum = 0.0;
    int bad = 0, skip = 0;
    const int row_end = it.row0 + it.rows;
    int step = 0;
    for (int base = it.row0; base < row_end; base += 64, ++step) {
        const int row = base + lane;
        bool live = row < row_end;
        const int rr = live ? row : it.row0;
        const double wb = R.w64 ? R.w64[rr] : 1.0;
        live = live && (wb > 0.0);
        fv8 xv;
        load_x(R.cols, R.n, R.v, rr, xv);
        Tape2<float, BIG> tape;
        const float pred = ft_forward<float, BIG, true>(fc, d.m, cst, xv, R.cols, R.n, rr, tape);
        double p64 = (double)pred;
        if (POLISH) {
            dv8 xd;
            load_x(R.cols64, R.n, R.v, rr, xd);
            Tape2<double, BIG> none;
            p64 = ft_forward<double, BIG, false>(fc, d.m, cp64 + (size_t)cand * FT_KMAX, xd,
                                                 R.cols64, R.n, rr, none);
        }
        const double y64 = R.y64[rr];
        const bool ok = live && isfinite(p64) && (polish || fabsf(pred) <= 1e30f) &&
                        isfinite(p64 - y64);
        const RcRowFit rf = rc_row_fit(wb, p64, y64, d.metric);
        const double r64 = rf.r, s64 = rf.s;
        const float seed = ok ? (float)s64 : 0.0f;
        const bool gfin = ft_reverse<float, BIG>(rc, d.nrev, seed, tape, X, lane, (float *)nullptr);
        const bool use = ok && gfin;
        if (polish) {
            if (live && !ok) bad = 1;
            if (ok && !gfin) ++skip;
            if (ok) lsum += wb * rf.rho;
        } else {
            if (live && !use) bad = 1;
            if (use) lsum += wb * rf.rho;
        }
        if (__any(ok && !gfin)) {
            if (!use)
                for (int j = 0; j < k; ++j) X[j * XS + lane] = 0.0f;
        }
        const double rs = s64 * r64;
        const float rhi = (float)rs;
        const float rlo = (float)(rs - (double)rhi);
        X[k * XS + lane] = use ? rhi : 0.0f;
        X[(k + 1) * XS + lane] = use ? rlo : 0.0f;
        __syncthreads();
        const int col = lane & 15;
        const int sub = lane >> 4;
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const float a0 = X[col * XS + 4 * q + sub];
            acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, a0, acc00, 0, 0, 0);
            if (KC == 32) {
                const float a1 = X[(16 + col) * XS + 4 * q + sub];
                acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, a0, acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, a1, acc11, 0, 0, 0);
            }
        }
        __syncthreads();
        if ((step & 3) == 3) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                g00[r] += (double)acc00[r];
                acc00[r] = 0.0f;
                if (KC == 32) {
                    g10[r] += (double)acc10[r];
                    g11[r] += (double)acc11[r];
                    acc10[r] = 0.0f;
                    acc11[r] = 0.0f;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        g00[r] += (double)acc00[r];
        if (KC == 32) {
            g10[r] += (double)acc10[r];
            g11[r] += (double)acc11[r];
        }
    }
    /* Lane l, register r holds D[4*(l/16) + r][l % 16] of each 16x16 block. */
    const int K = k + 2;
    double *pg = pgram + (size_t)item * (size_t)tri_stride;
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * (lane >> 4) + r;
        if (i >= j && i < K) pg[i * (i + 1) / 2 + j] = g00[r];
        if (KC == 32) {
            const int i1 = 16 + i;
            if (i1 < K) pg[i1 * (i1 + 1) / 2 + j] = g10[r];
            if (i >= j && i1 < K) pg[i1 * (i1 + 1) / 2 + 16 + j] = g11[r];
        }
    }
    const double ls = rc_wave_sum_f64(lsum);
    int sk = skip;
    for (int off = 32; off > 0; off >>= 1) sk += __shfl_xor(sk, off, 64);
    const int anybad = __any(bad) ? 1 : 0;
    if (lane == 0) {
        ploss[item] = ls;
        pbad[item] = anybad;
        pskip[item] = sk;
    }
}

/* The plain FP64 arm's assembly on the general interpreter: ft_asm_kernel's
   row contract with f64 rows, f64 constants, the f64 value and adjoint
   sweeps, the augmented row [s*J | s*r | 0] in an f64 LDS block and the Gram
   on FP64 MFMA (v_mfma_f64_16x16x4f64), accumulated in f64.  A row is usable
   when its f64 prediction and residual are finite. */
constexpr int XS64 = 68; /* f64 column stride: 64 rows + 4 */
template <int KC, bool BIG, bool POLISH>
__global__ __launch_bounds__(64) void ft_asm64_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run,
    const double *__restrict__ cp64, double *__restrict__ pgram, int tri_stride,
    double *__restrict__ ploss, int *__restrict__ pbad, int *__restrict__ pskip)
{
    typedef double double4v __attribute__((ext_vector_type(4)));
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const FtCand d = cands[cand];
    const FtRows R = rowsd[d.rowset];
    const int lane = threadIdx.x;
    const int k = d.k;

    __shared__ double X[KC * XS64];
    for (int i = lane; i < KC * XS64; i += 64) X[i] = 0.0;
    __syncthreads();

    const double *cst = cp64 + (size_t)cand * FT_KMAX;
    const uint32_t *fc = code + d.fwd_off;
    const uint32_t *rc = code + d.rev_off;
    double4v acc00 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0}, acc11 = {0, 0, 0, 0};
    double lsum = 0.0;
    int bad = 0, skip = 0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64) {
        const int row = base + lane;
        bool live = row < row_end;
        const int rr = live ? row : it.row0;
        const double wb = R.w64 ? R.w64[rr] : 1.0;
        live = live && (wb > 0.0);
        dv8 xd;
        load_x(R.cols64, R.n, R.v, rr, xd);
        Tape2<double, BIG> tape;
        const double pred = ft_forward<double, BIG, true>(fc, d.m, cst, xd, R.cols64, R.n, rr, tape);
        const double y64 = R.y64[rr];
        const bool ok = live && isfinite(pred) && isfinite(pred - y64);
        const RcRowFit rf = rc_row_fit(wb, pred, y64, d.metric);
        const double seed = ok ? rf.s : 0.0;
        const bool gfin = ft_reverse<double, BIG, double, XS64>(rc, d.nrev, seed, tape, X, lane,
                                                                (double *)nullptr);
        const bool use = ok && gfin;
        if (POLISH) {
            if (live && !ok) bad = 1;
            if (ok && !gfin) ++skip;
            if (ok) lsum += wb * rf.rho;
        } else {
            if (live && !use) bad = 1;
            if (use) lsum += wb * rf.rho;
        }
        if (__any(ok && !gfin)) {
            if (!use)
                for (int j = 0; j < k; ++j) X[j * XS64 + lane] = 0.0;
        }
        X[k * XS64 + lane] = use ? rf.s * rf.r : 0.0;
        X[(k + 1) * XS64 + lane] = 0.0;
        __syncthreads();
        const int col = lane & 15;
        const int sub = lane >> 4;
#pragma unroll 4
        for (int q = 0; q < 16; ++q) {
            const double a0 = X[col * XS64 + 4 * q + sub];
            acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc00, 0, 0, 0);
            if (KC == 32) {
                const double a1 = X[(16 + col) * XS64 + 4 * q + sub];
                acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a0, acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, a1, acc11, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    /* FP64 MFMA output: lane l, register r holds D[l/16 + 4 r][l % 16]. */
    const int K = k + 2;
    double *pg = pgram + (size_t)item * (size_t)tri_stride;
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = (lane >> 4) + 4 * r;
        if (i >= j && i < K) pg[i * (i + 1) / 2 + j] = acc00[r];
        if (KC == 32) {
            const int i1 = 16 + i;
            if (i1 < K) pg[i1 * (i1 + 1) / 2 + j] = acc10[r];
            if (i >= j && i1 < K) pg[i1 * (i1 + 1) / 2 + 16 + j] = acc11[r];
        }
    }
    const double ls = rc_wave_sum_f64(lsum);
    int sk = skip;
    for (int off = 32; off > 0; off >>= 1) sk += __shfl_xor(sk, off, 64);
    const int anybad = __any(bad) ? 1 : 0;
    if (lane == 0) {
        ploss[item] = ls;
        pbad[item] = anybad;
        pskip[item] = sk;
    }
}

/* ---- trial / exact loss: forward only ------------------------------------ */

template <typename T, bool BIG>
__global__ __launch_bounds__(64) void ft_trial_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run, const T *__restrict__ cp,
    double *__restrict__ ploss)
{
    /* Rows per lane: 4 for f32 and 2 for f64 in the common class keep the
       stack in VGPRs; the oversized class runs one row per lane. */
    constexpr int R = BIG ? 1 : (sizeof(T) == 4 ? 4 : 2);
    constexpr int S = BIG ? BIG_S : SMALL_S;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const FtCand d = cands[cand];
    const FtRows R_ = rowsd[d.rowset];
    const int lane = threadIdx.x;
    const T *cst = cp + (size_t)cand * FT_KMAX;
    const uint32_t *fc = code + d.t_off;
    const T *cols;
    if (sizeof(T) == 4) cols = (const T *)(const void *)R_.cols;
    else cols = (const T *)(const void *)R_.cols64;
    double lsum = 0.0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64 * R) {
        int xrow[R];
        bool live[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + r * 64 + lane;
            live[r] = row < row_end;
            xrow[r] = live[r] ? row : it.row0;
        }
        T pred[R];
        ft_value_rows<T, S, R>(fc, d.t_m, cst, cols, R_.n, xrow, pred);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double wb = R_.w64 ? R_.w64[xrow[r]] : 1.0;
            if (!live[r] || !(wb > 0.0)) continue;
            const double yr = R_.y64[xrow[r]];
            const double rr = (double)pred[r] - yr;
            bool hard = !isfinite(pred[r]) || !isfinite(rr);
            if (sizeof(T) == 4) hard = hard || fabs((double)pred[r]) > 1e30;
            lsum += hard ? 1e30 : wb * rc_row_stat((double)pred[r], yr, d.metric);
        }
    }
    const double ls = rc_wave_sum_f64(lsum);
    if (lane == 0) ploss[item] = ls;
}

/* ---- per-row probe (test hook): pred, constant and input gradients -------- */

__global__ __launch_bounds__(64) void ft_probe_kernel(const FtCand *__restrict__ cands,
                                                      const FtRows *__restrict__ rowsd,
                                                      const uint32_t *__restrict__ code,
                                                      const float *__restrict__ cp32, int cand,
                                                      int rows, float *__restrict__ out_pred,
                                                      float *__restrict__ out_jac, int kmax,
                                                      float *__restrict__ out_gx)
{
    __shared__ float X[32 * XS];
    const int lane = threadIdx.x;
    const FtCand d = cands[cand];
    const FtRows R = rowsd[d.rowset];
    const int row = blockIdx.x * 64 + lane;
    const bool live = row < rows && row < R.n;
    const int rr = live ? row : 0;
    for (int i = lane; i < 32 * XS; i += 64) X[i] = 0.0f;
    __syncthreads();
    fv8 xv;
    load_x(R.cols, R.n, R.v, rr, xv);
    Tape2<float, true> tape;
    const float pred = ft_forward<float, true, true>(code + d.fwd_off, d.m, cp32 + (size_t)cand * FT_KMAX,
                                                     xv, R.cols, R.n, rr, tape);
    float gx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ft_reverse<float, true>(code + d.rev_off, d.nrev, 1.0f, tape, X, lane, gx);
    __syncthreads();
    if (!live) return;
    out_pred[row] = pred;
    for (int j = 0; j < d.k; ++j) out_jac[(size_t)row * kmax + j] = X[j * XS + lane];
    for (int v = 0; v < R.v && v < 8; ++v) out_gx[(size_t)row * R.v + v] = gx[v];
}

/* ---- control: one wave per candidate --------------------------------------- */

__device__ __forceinline__ double clamp_step(double t) { return fmin(1e6, fmax(-1e6, t)); }

/* A solve the serial control step leaves to the whole wave: the LM step at
   the cached normal equations or the polish's Gauss-Newton step. */
enum { CT_SOLVE_NONE = 0, CT_SOLVE_LM = 1, CT_SOLVE_POLISH = 2 };

__device__ __forceinline__ int ctrl_lane0(int op, int cand, FtState *s, const FtCand &d,
                                       const double *__restrict__ c0, int kmax,
                                       float *__restrict__ cp32, double *__restrict__ cp64,
                                       const double *G, double sh_loss, int sh_bad, int sh_skip,
                                       int ran_trial, bool want_gram, double &solve_lambda,
                                       double *__restrict__ outc, double *__restrict__ outsse,
                                       int *__restrict__ outstatus, int all64);
__device__ __forceinline__ void ctrl_after_solve(int kind, int rc, int cand, FtState *s, int k,
                                              const double *delta, float *__restrict__ cp32,
                                              double *__restrict__ cp64, int all64);
__device__ __forceinline__ void ctrl_flags(int op, int cand, const FtState *s,
                                        int *__restrict__ run_asm, int *__restrict__ run_trial);

__global__ __launch_bounds__(64) void ft_ctrl_kernel(
    int op, int ncand, FtState *__restrict__ S, const FtCand *__restrict__ cands,
    const double *__restrict__ c0, int kmax, int *__restrict__ run_asm, int *__restrict__ run_trial,
    float *__restrict__ cp32, double *__restrict__ cp64, const double *__restrict__ pgram,
    int tri_stride, const double *__restrict__ ploss, const int *__restrict__ pbad,
    const int *__restrict__ pskip, double *__restrict__ outc, double *__restrict__ outsse,
    int *__restrict__ outstatus, int all64)
{
    const int cand = blockIdx.x;
    if (cand >= ncand) return;
    const int lane = threadIdx.x;
    const FtCand d = cands[cand];
    __shared__ double G[FT_TRI_MAX];
    __shared__ double L[FT_KMAX * FT_KMAX];
    __shared__ double sh_loss;
    __shared__ int sh_bad, sh_skip;
    /* The candidate's state is staged in LDS for this launch: the serial
       control below (lane 0) and its solve then read and write LDS instead
       of global memory. */
    __shared__ __align__(16) FtState sS;
    constexpr int SW = (int)(sizeof(FtState) / sizeof(int4));
    constexpr int SU = (SW + 63) / 64;
    {
        const int4 *src = (const int4 *)(S + cand);
        int4 *dst = (int4 *)&sS;
        /* All loads first (lanes past the end repeat the last word, which
           they also store, so no branch splits the loads from each other). */
        int4 v[SU];
#pragma unroll
        for (int u = 0; u < SU; ++u) v[u] = src[min(lane + 64 * u, SW - 1)];
#pragma unroll
        for (int u = 0; u < SU; ++u) dst[min(lane + 64 * u, SW - 1)] = v[u];
    }
    __syncthreads();
    FtState *s = &sS;

    const int ran_asm = run_asm[cand];
    const int ran_trial = run_trial[cand];
    const int k = d.k;
    const int K = k + 2;
    const int tri = K * (K + 1) / 2;

    /* Fixed-order reductions over this candidate's items. */
    const bool want_gram = (op == CT_AFTER_ASM || op == CT_AFTER_ASM_P) && ran_asm &&
                           (s->stage == ST_NEED_ASM || s->stage == ST_P_NEED_ASM);
    const bool want_loss = ((op == CT_AFTER_TRIAL || op == CT_AFTER_TRIAL_P || op == CT_REBASE) &&
                            ran_trial) || want_gram;
    if (want_gram) {
        for (int q = lane; q < tri; q += 64)
            G[q] = rc_ordered_sum(pgram + (size_t)d.item0 * (size_t)tri_stride + q,
                                  (size_t)tri_stride, d.nitems);
    }
    if (lane == 0 && want_gram) s->n_asm += 1;
    if (lane == 0 && want_loss && !want_gram) {
        if (op == CT_AFTER_TRIAL) s->n_trial += 1;
        else s->n_exact += 1;
    }
    if (want_loss && lane == 0) {
        int b = 0, sk = 0;
        if (want_gram) {
            for (int t0 = 0; t0 < d.nitems; t0 += 8) {
                int vb[8], vs[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int t = d.item0 + min(t0 + u, d.nitems - 1);
                    vb[u] = pbad[t];
                    vs[u] = pskip[t];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (t0 + u < d.nitems) {
                        b |= vb[u];
                        sk += vs[u];
                    }
            }
        }
        sh_loss = rc_ordered_sum(ploss + d.item0, 1, d.nitems);
        sh_bad = b;
        sh_skip = sk;
    }
    __shared__ int sh_solve;
    __shared__ double sh_lambda;
    __shared__ double sh_delta[FT_KMAX], sh_dsc[FT_KMAX], sh_rhs[FT_KMAX];
    __syncthreads();
    if (lane == 0) {
        double lam = 0.0;
        sh_solve = ctrl_lane0(op, cand, s, d, c0, kmax, cp32, cp64, G, sh_loss, sh_bad, sh_skip,
                              ran_trial, want_gram, lam, outc, outsse, outstatus, all64);
        sh_lambda = lam;
    }
    __syncthreads();
    const int solve = sh_solve;
    if (solve != CT_SOLVE_NONE) {
        const int rc = ft_solve_wave(s->jtj, s->jtr, sh_lambda, k, sh_delta, L, sh_dsc, sh_rhs,
                                     lane);
        if (lane == 0) ctrl_after_solve(solve, rc, cand, s, k, sh_delta, cp32, cp64, all64);
    }
    if (lane == 0) ctrl_flags(op, cand, s, run_asm, run_trial);
    __syncthreads();
    {
        const int4 *src = (const int4 *)&sS;
        int4 *dst = (int4 *)(S + cand);
        for (int i = lane; i < SW; i += 64) dst[i] = src[i];
    }
}

/* Staging of a pass's constants: f32 for the assembly / trial passes, f64
   for the exact and polish passes (zero beyond k). */
__device__ __forceinline__ void ctrl_set_pass32(float *__restrict__ cp32, int cand, int k,
                                             const double *v)
{
    for (int j = 0; j < FT_KMAX; ++j) cp32[(size_t)cand * FT_KMAX + j] = j < k ? (float)v[j] : 0.0f;
}
__device__ __forceinline__ void ctrl_set_pass64(double *__restrict__ cp64, int cand, int k,
                                             const double *v)
{
    for (int j = 0; j < FT_KMAX; ++j) cp64[(size_t)cand * FT_KMAX + j] = j < k ? v[j] : 0.0;
}

/* The serial part of a control step (one lane): LM / IRLS / polish state
   transitions on the candidate's LDS-staged state.  A step that needs a
   solve stops there and returns its kind (with its lambda); the wave then
   solves and ctrl_after_solve finishes the step. */
__device__ __forceinline__ int ctrl_lane0(int op, int cand, FtState *s, const FtCand &d,
                                       const double *__restrict__ c0, int kmax,
                                       float *__restrict__ cp32, double *__restrict__ cp64,
                                       const double *G, double sh_loss, int sh_bad, int sh_skip,
                                       int ran_trial, bool want_gram, double &solve_lambda,
                                       double *__restrict__ outc, double *__restrict__ outsse,
                                       int *__restrict__ outstatus, int all64)
{
    const int k = d.k;
    double *cc = s->c;
    int solve = CT_SOLVE_NONE;
    /* The LM loop's passes read f32 constants, or f64 ones in the plain
       FP64 arm (every pass on f64 rows). */
    auto set_pass32 = [&](const double *v) {
        if (all64) ctrl_set_pass64(cp64, cand, k, v);
        else ctrl_set_pass32(cp32, cand, k, v);
    };
    auto set_pass64 = [&](const double *v) { ctrl_set_pass64(cp64, cand, k, v); };
    auto take_gram = [&]() {
        for (int i = 0; i < k; ++i) {
            for (int j = 0; j <= i; ++j) s->jtj[i * (i + 1) / 2 + j] = G[i * (i + 1) / 2 + j];
            s->jtr[i] = G[k * (k + 1) / 2 + i] + G[(k + 1) * (k + 2) / 2 + i];
        }
    };
    auto lm_next = [&]() {
        /* After an iteration: start the next one or end the loop. */
        if (s->stage == ST_LM_DONE) return;
        if (s->assembled) {
            /* Solve at the cached normal equations (the wave's solve; the
               trial is staged in ctrl_after_solve). */
            solve = CT_SOLVE_LM;
            solve_lambda = s->lambda;
        } else {
            set_pass32(cc);
            s->stage = ST_NEED_ASM;
        }
    };
    auto end_inner = [&]() {
        /* The inner loop of this round ended (count or break). */
        s->round += 1;
        s->it = 0;
        if (s->round >= s->rounds || s->best == 0.0 || s->status != 0) {
            s->stage = ST_LM_DONE;
        } else {
            s->assembled = 0; /* new IRLS weights, new Gram */
        }
    };

    switch (op) {
        case CT_INIT: {
            s->k = k;
            s->metric = d.metric;
            s->irls = (d.metric == RC_LOSS_ABSOLUTE);
            const int it = s->iters;
            s->rounds = s->irls ? (it < 1 ? 1 : (it < 4 ? it : 4)) : 1;
            s->inner = s->irls ? (it / s->rounds < 1 ? 1 : it / s->rounds) : it;
            for (int j = 0; j < FT_KMAX; ++j) cc[j] = j < k ? c0[(size_t)cand * kmax + j] : 0.0;
            s->lambda = 1e-3;
            s->best = 0.0;
            s->status = 0;
            s->sstart = 0;
            s->srank = 0;
            s->improved = 0;
            s->skipped = 0;
            s->it = 0;
            s->round = 0;
            s->assembled = 0;
            s->first = 1;
            s->halvings = 0;
            s->polish_it = 0;
            s->n_asm = 0;
            s->n_trial = 0;
            s->n_exact = 0;
            s->acc_bits = 0;
            s->n_decided = 0;
            if (s->pre_status >= 0) {
                s->stage = ST_DONE;
            } else if (s->iters > 0) {
                set_pass32(cc);
                s->stage = ST_NEED_ASM;
            } else {
                s->stage = ST_LM_DONE;
            }
            break;
        }
        case CT_AFTER_ASM: {
            if (!want_gram) break;
            const double loss = sh_loss;
            if (s->first) {
                s->first = 0;
                s->best = loss;
                s->sstart = sh_bad ? 0 : 1;
                if (sh_bad) { s->status = 2; s->stage = ST_LM_DONE; break; }
                if (s->best == 0.0) { s->stage = ST_LM_DONE; break; }
            } else if (sh_bad) {
                s->status = 2;
                s->stage = ST_LM_DONE;
                break;
            }
            take_gram();
            s->assembled = 1;
            lm_next();
            break;
        }
        case CT_AFTER_TRIAL: {
            if (!ran_trial || s->stage != ST_NEED_TRIAL) break;
            const double tsse = s->pruned ? __builtin_inf() : sh_loss;
            s->pruned = 0;
            if (!(tsse < s->best) && s->halvings > 0) {
                /* True-metric Armijo halving for the L1 arm (eval.cu:2496-2509). */
                s->halvings -= 1;
                for (int j = 0; j < k; ++j) s->trial[j] = cc[j] + 0.5 * (s->trial[j] - cc[j]);
                set_pass32(s->trial);
                break;
            }
            bool brk = false;
            if (tsse < s->best && s->n_decided < 31) s->acc_bits |= 1 << s->n_decided;
            s->n_decided += 1;
            if (tsse < s->best) {
                for (int j = 0; j < k; ++j) cc[j] = s->trial[j];
                s->best = tsse;
                s->lambda = fmax(s->lambda / 8.0, 1e-12);
                s->improved = 1;
                s->assembled = 0;
                if (s->best == 0.0) brk = true;
            } else {
                s->lambda *= 6.0;
                if (s->lambda > 1e9) brk = true;
            }
            s->it += 1;
            if (brk || s->it >= s->inner) end_inner();
            lm_next();
            break;
        }
        case CT_PRE_REBASE: {
            if (s->stage != ST_DONE) s->stage = ST_LM_DONE;
            if (s->stage == ST_LM_DONE && s->status == 0) set_pass64(cc);
            break;
        }
        case CT_REBASE: {
            if (s->stage != ST_LM_DONE) break;
            if (ran_trial && s->status == 0 && isfinite(sh_loss)) s->best = sh_loss;
            if (s->iters > 0 && s->status == 0 && s->best != 0.0) {
                s->polish_it = 0;
                set_pass32(cc);
                set_pass64(cc);
                s->stage = ST_P_NEED_ASM;
            } else {
                s->stage = ST_DONE;
            }
            break;
        }
        case CT_AFTER_ASM_P: {
            if (!want_gram) break;
            if (sh_bad) { s->status = 2; s->stage = ST_DONE; break; }
            s->skipped += sh_skip;
            take_gram();
            solve = CT_SOLVE_POLISH;
            solve_lambda = 0.0;
            break;
        }
        case CT_AFTER_TRIAL_P: {
            if (!ran_trial || s->stage != ST_P_NEED_TRIAL) break;
            const double tsse = sh_loss;
            if (!(tsse < s->best) && s->halvings > 0) {
                s->halvings -= 1;
                for (int j = 0; j < k; ++j) s->trial[j] = cc[j] + 0.5 * (s->trial[j] - cc[j]);
                set_pass64(s->trial);
                break;
            }
            if (tsse < s->best) {
                for (int j = 0; j < k; ++j) cc[j] = s->trial[j];
                s->best = tsse;
                s->improved = 1;
                s->polish_it += 1;
                if (s->polish_it < 4 && s->best != 0.0) {
                    set_pass32(cc);
                    set_pass64(cc);
                    s->stage = ST_P_NEED_ASM;
                } else {
                    s->stage = ST_DONE;
                }
            } else {
                s->stage = ST_DONE;
            }
            break;
        }
        case CT_FINAL: {
            int st = s->status;
            if (s->pre_status >= 0) {
                st = s->pre_status;
            } else if (st == 1 || st == 2) {
                st = !s->sstart ? 7 : (s->improved ? 5 : 6);
            } else if (st == 0) {
                if (s->skipped > 0) st = 8;
                else if (s->srank) st = 4;
                else if (!s->improved && s->iters > 0) st = 6;
            }
            for (int j = 0; j < kmax; ++j)
                outc[(size_t)cand * kmax + j] = j < k ? cc[j] : c0[(size_t)cand * kmax + j];
            outsse[cand] = s->pre_status >= 0 ? (s->pre_status == 7 ? 0.0 : __builtin_nan(""))
                                              : s->best;
            outstatus[cand] = st;
            break;
        }
        default: break;
    }
    return solve;
}

/* The end of a control step that solved: stage the trial (clamped step from
   the current constants) or record the failed solve. */
__device__ __forceinline__ void ctrl_after_solve(int kind, int rc, int cand, FtState *s, int k,
                                              const double *delta, float *__restrict__ cp32,
                                              double *__restrict__ cp64, int all64)
{
    if (rc == 1) {
        if (kind == CT_SOLVE_LM) {
            s->status = 1;
            s->stage = ST_LM_DONE;
        } else {
            s->stage = ST_DONE;
        }
        return;
    }
    if (rc == 4) s->srank = 1;
    for (int j = 0; j < k; ++j) s->trial[j] = clamp_step(s->c[j] + delta[j]);
    if (kind == CT_SOLVE_LM) {
        if (all64) ctrl_set_pass64(cp64, cand, k, s->trial);
        else ctrl_set_pass32(cp32, cand, k, s->trial);
        s->halvings = s->irls ? 2 : 0;
        s->stage = ST_NEED_TRIAL;
    } else {
        ctrl_set_pass64(cp64, cand, k, s->trial);
        s->halvings = s->irls ? 1 : 0;
        s->stage = ST_P_NEED_TRIAL;
    }
}

/* Flags for the next pass, from the step's final stage. */
__device__ __forceinline__ void ctrl_flags(int op, int cand, const FtState *s,
                                        int *__restrict__ run_asm, int *__restrict__ run_trial)
{
    const int stg = s->stage;
    switch (op) {
        case CT_INIT:
        case CT_AFTER_ASM:
        case CT_AFTER_TRIAL:
            run_asm[cand] = (stg == ST_NEED_ASM);
            run_trial[cand] = (stg == ST_NEED_TRIAL);
            break;
        case CT_PRE_REBASE:
            run_asm[cand] = 0;
            run_trial[cand] = (stg == ST_LM_DONE && s->status == 0);
            break;
        case CT_REBASE:
        case CT_AFTER_ASM_P:
        case CT_AFTER_TRIAL_P:
            run_asm[cand] = (stg == ST_P_NEED_ASM);
            run_trial[cand] = (stg == ST_P_NEED_TRIAL);
            break;
        default:
            run_asm[cand] = 0;
            run_trial[cand] = 0;
            break;
    }
}

__global__ void ft_state_init(FtState *S, int ncand, const int *pre_status, int iters)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    S[c].pre_status = pre_status[c];
    S[c].iters = iters;
    S[c].pruned = 0;
}

/* Between the two halves of an LM trial pass.  A candidate's trial loss is
   the sum of its items' losses in item order (the control kernel's
   reduction) and every term is non-negative, so once the sum of its first
   (nitems + 1) / 2 items, taken in the same order, reaches the best loss,
   the whole sum does too and the trial is rejected whatever the rest adds:
   the second half of its items is skipped (run_b = 0) and the control
   kernel reads the trial as rejected.  The accept / reject sequence and
   every accepted loss are those of the full pass. */
__global__ void ft_trial_prune(int nprog, FtState *__restrict__ S, const FtCand *__restrict__ cands,
                               const double *__restrict__ ploss, const int *__restrict__ run_t,
                               int *__restrict__ run_b)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nprog) return;
    FtState *st = S + c;
    int pr = 0;
    if (run_t[c] && st->stage == ST_NEED_TRIAL && isfinite(st->best)) {
        const FtCand d = cands[c];
        const int h = (d.nitems + 1) / 2;
        pr = rc_ordered_sum(ploss + d.item0, 1, h) >= st->best ? 1 : 0;
    }
    st->pruned = pr;
    run_b[c] = (run_t[c] && !pr) ? 1 : 0;
}

} /* namespace */

/* ======================================================================== */
/* Host side.                                                                */

/* Operators with an analytic value/adjoint pair (fit.rs:890-920). */
static bool fit_op_ok(uint32_t op)
{
    switch (op) {
        case OP_CONST: case OP_VAR: case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV:
        case OP_SIN: case OP_COS: case OP_TAN: case OP_EXP: case OP_LOG: case OP_SQRT:
        case OP_ABS: case OP_TANH: case OP_SIG: case OP_SQUARE: case OP_CUBE: case OP_NEG:
        case OP_GAUSS: case OP_POW: case OP_SINH: case OP_COSH: case OP_ERF: case OP_ERFC:
            return true;
        default: return false;
    }
}

static bool is_bin(uint32_t op)
{
    return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV || op == OP_POW;
}

/* Per-candidate translation result. */
struct FtProg {
    std::vector<uint32_t> fwd, rev;
    int pre_status = -1; /* >= 0: decided on the host (no device work) */
    int max_stack = 0, max_adj = 0, tape = 0, tape_l = 0, tape_r = 0;
};

/* Translate one candidate's records.  want_const / want_var select which
   leaves the reverse sweep must reach. */
static FtProg translate_fit(const uint32_t *w, int words, int p, int nvars, bool want_const,
                            bool want_var)
{
    FtProg out;
    const int m = words / 4;
    if (words % 4 != 0 || m < 1 || m > FT_MAX_NODES || p < 0 || p > FT_KMAX) {
        out.pre_status = 1;
        return out;
    }
    std::vector<int> st;
    st.reserve((size_t)m);
    std::vector<char> onpath(m, 0);
    bool unsupported = false;
    for (int i = 0; i < m; ++i) {
        const uint32_t op = w[4 * i], k0 = w[4 * i + 1], k1 = w[4 * i + 2], slot = w[4 * i + 3];
        if (!fit_op_ok(op)) unsupported = true;
        if (op == OP_CONST) {
            if ((int)slot >= p) { out.pre_status = 1; return out; }
            onpath[i] = want_const;
            st.push_back(i);
        } else if (op == OP_VAR) {
            if ((int)slot >= nvars || slot > 0xffffu) { out.pre_status = 1; return out; }
            onpath[i] = want_var;
            st.push_back(i);
        } else if (is_bin(op)) {
            if (st.size() < 2 || st.back() != (int)k1 || st[st.size() - 2] != (int)k0) {
                out.pre_status = 1; /* not the canonical post-order tree */
                return out;
            }
            st.pop_back();
            st.pop_back();
            onpath[i] = onpath[k0] || onpath[k1];
            st.push_back(i);
        } else {
            if (st.empty() || st.back() != (int)k0) { out.pre_status = 1; return out; }
            st.pop_back();
            onpath[i] = onpath[k0];
            st.push_back(i);
        }
    }
    if (st.size() != 1 || st[0] != m - 1) { out.pre_status = 1; return out; }
    if (unsupported) {
        /* The reference device forward yields NaN for every row: a non-finite
           start, status 7, constants unchanged. */
        out.pre_status = 7;
        return out;
    }
    /* Forward stream, stack depth and tape size. */
    int depth = 0, tape = 0;
    out.fwd.reserve((size_t)m);
    out.rev.reserve((size_t)m);
    for (int i = 0; i < m; ++i) {
        const uint32_t op = w[4 * i], k0 = w[4 * i + 1], k1 = w[4 * i + 2], slot = w[4 * i + 3];
        uint32_t fl = 0, arg = 0;
        if (op == OP_CONST || op == OP_VAR) {
            arg = slot;
            ++depth;
        } else if (is_bin(op)) {
            if (onpath[k0]) fl |= FF_D0;
            if (onpath[k1]) fl |= FF_D1;
            /* ADD/SUB partials are the constants +-1: no tape. */
            if (op == OP_ADD || op == OP_SUB) fl = 0;
            --depth;
        } else {
            if (onpath[k0] && op != OP_NEG) fl |= FF_D0;
        }
        tape += ((fl & FF_D0) ? 1 : 0) + ((fl & FF_D1) ? 1 : 0);
        out.tape_l += (fl & FF_D0) ? 1 : 0;
        out.tape_r += (fl & FF_D1) ? 1 : 0;
        out.max_stack = std::max(out.max_stack, depth);
        out.fwd.push_back(op | (fl << 8) | (arg << 16));
    }
    out.tape = tape;
    /* Reverse stream over on-path nodes in reverse post-order. */
    int ad = 1, adpeak = 1;
    for (int i = m - 1; i >= 0; --i) {
        if (!onpath[i]) continue;
        const uint32_t op = w[4 * i], k0 = w[4 * i + 1], k1 = w[4 * i + 2], slot = w[4 * i + 3];
        uint32_t kind, fl = 0, arg = 0;
        --ad;
        if (op == OP_CONST) { kind = RK_CONST; arg = slot; }
        else if (op == OP_VAR) { kind = RK_VAR; arg = slot; }
        else if (is_bin(op)) {
            kind = (op == OP_ADD) ? RK_ADD : (op == OP_SUB) ? RK_SUB : RK_BINARY;
            if (onpath[k0]) { fl |= RF_LEFT; ++ad; }
            if (onpath[k1]) { fl |= RF_RIGHT; ++ad; }
        } else {
            kind = (op == OP_NEG) ? RK_NEG : RK_UNARY;
            ++ad;
        }
        adpeak = std::max(adpeak, ad);
        out.rev.push_back(kind | (fl << 8) | (arg << 16));
    }
    out.max_adj = adpeak;
    return out;
}

/* Device-resident state of the fit arm (grown at the high-water mark). */
struct RcFitState {
    unsigned char *dev = nullptr;
    size_t dev_bytes = 0;
    unsigned char *host = nullptr;
    size_t host_bytes = 0;
    int mode = 0; /* 0 mixed, 1 plain f64 */
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<uint64_t> graph_key;
    /* Profiling (rc_fit_profile): per-pass event timing, launches eagerly. */
    int profile = 0;
    double t_pass[4] = {0, 0, 0, 0};  /* asm, trial32, exact64, ctrl */
    int n_pass[4] = {0, 0, 0, 0};
    FtState *last_state = nullptr;
    int last_nprog = 0;
    /* A second stream with its fork / join events: the small class
       kernels of a pass run on it beside the large ones. */
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
};

void rc_fit_state_free(RcCtx *c)
{
    if (!c->fit) return;
    if (c->fit->exec) (void)hipGraphExecDestroy(c->fit->exec);
    if (c->fit->graph) (void)hipGraphDestroy(c->fit->graph);
    if (c->fit->dev) (void)hipFree(c->fit->dev);
    if (c->fit->host) (void)hipHostFree(c->fit->host);
    if (c->fit->side) (void)hipStreamDestroy(c->fit->side);
    if (c->fit->ev_fork) (void)hipEventDestroy(c->fit->ev_fork);
    if (c->fit->ev_join) (void)hipEventDestroy(c->fit->ev_join);
    delete c->fit;
    c->fit = nullptr;
}

int rc_rows_ensure_f64(RcCtx *c, RcRows *r);
extern "C" int rch_fitvm_compile(const unsigned int *rec, int words, int p, int nvars, int col_bytes,
                                 unsigned int flags, unsigned int *fwd, int cap_f,
                                 unsigned int *rev, int cap_r, int *info);
static size_t take(size_t &cur, size_t bytes)
{
    size_t o = (cur + 255) & ~(size_t)255;
    cur = o + bytes;
    return o;
}

/* Byte offsets of the fit region: inputs [code | cands | rows | items |
   lists | pre | c0], outputs [c | sse | status], then work.  The recorded
   graph bakes every one of them into its kernel arguments, so the graph
   key holds them all. */
struct FtRegion {
    size_t o_code, o_cand, o_rows, o_items, o_lists, o_pre, o_c0, in_end;
    size_t o_outc, o_sse, o_st, io_end;
    size_t o_state, o_run_a, o_run_t, o_run_b, o_cp32, o_cp64, o_pgram, o_ploss, o_pbad, o_pskip,
        o_p64, dev_end;
};
static const int FT_REGION_FIELDS = (int)(sizeof(FtRegion) / sizeof(size_t));

static FtRegion ft_region(size_t code_words, size_t ncands, size_t nrowsets, size_t nitems,
                          size_t nlist_words, int nprog, int kmax, int tri_stride,
                          size_t p64_bytes = 0)
{
    FtRegion g;
    size_t cur = 0;
    const size_t np = (size_t)nprog, ni = std::max<size_t>(1, nitems);
    g.o_code = take(cur, code_words * 4);
    g.o_cand = take(cur, ncands * sizeof(FtCand));
    g.o_rows = take(cur, nrowsets * sizeof(FtRows));
    g.o_items = take(cur, ni * sizeof(FtItem));
    g.o_lists = take(cur, nlist_words * 4);
    g.o_pre = take(cur, np * 4);
    g.o_c0 = take(cur, np * (size_t)kmax * 8);
    g.in_end = take(cur, 0);
    g.o_outc = take(cur, np * (size_t)kmax * 8);
    g.o_sse = take(cur, np * 8);
    g.o_st = take(cur, np * 4);
    g.io_end = take(cur, 0);
    g.o_state = take(cur, np * sizeof(FtState));
    g.o_run_a = take(cur, np * 4);
    g.o_run_t = take(cur, np * 4);
    g.o_run_b = take(cur, np * 4);
    g.o_cp32 = take(cur, np * FT_KMAX * 4);
    g.o_cp64 = take(cur, np * FT_KMAX * 8);
    g.o_pgram = take(cur, ni * (size_t)tri_stride * 8);
    g.o_ploss = take(cur, ni * 8);
    g.o_pbad = take(cur, ni * 4);
    g.o_pskip = take(cur, ni * 4);
    g.o_p64 = take(cur, p64_bytes);
    g.dev_end = take(cur, 0);
    return g;
}

/* The replay key of a recorded fit graph: the batch shape, the region base
   and every region offset, then the class list offsets. */
static std::vector<uint64_t> ft_graph_key(int nprog, int nitems, int iters, int any_irls, int kmax,
                                          int tri_stride, const void *base, const FtRegion &g,
                                          const int *lofs, int nl, const int *tofs, int nt,
                                          int mode)
{
    std::vector<uint64_t> key = {(uint64_t)nprog, (uint64_t)nitems, (uint64_t)iters,
                                 (uint64_t)any_irls | ((uint64_t)mode << 8), (uint64_t)kmax,
                                 (uint64_t)tri_stride, (uint64_t)(uintptr_t)base};
    const size_t *fld = (const size_t *)&g;
    for (int q = 0; q < FT_REGION_FIELDS; ++q) key.push_back((uint64_t)fld[q]);
    for (int q = 0; q < nl; ++q) key.push_back((uint64_t)lofs[q]);
    for (int q = 0; q < nt; ++q) key.push_back((uint64_t)tofs[q]);
    return key;
}

static int fit_reserve(RcCtx *c, size_t dev_bytes, size_t host_bytes)
{
    RcFitState *f = c->fit;
    if (dev_bytes > f->dev_bytes) {
        if (f->dev) { (void)hipFree(f->dev); c->fs_alloc_calls++; if (c->fs_batches) c->fs_steady_alloc_calls++; }
        f->dev = nullptr;
        size_t want = dev_bytes + dev_bytes / 4;
        if (hipMalloc((void **)&f->dev, want) != hipSuccess) {
            (void)hipGetLastError();
            f->dev = nullptr;
            f->dev_bytes = 0;
            return -1;
        }
        f->dev_bytes = want;
        c->fs_alloc_calls++;
        c->fs_growths++;
        if (c->fs_batches) c->fs_steady_alloc_calls++;
        /* A new region invalidates the recorded launch sequence. */
        if (f->exec) { (void)hipGraphExecDestroy(f->exec); f->exec = nullptr; }
        if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
        f->graph_key.clear();
    }
    if (host_bytes > f->host_bytes) {
        if (f->host) { (void)hipHostFree(f->host); c->fs_alloc_calls++; }
        f->host = nullptr;
        size_t want = host_bytes + host_bytes / 4;
        if (hipHostMalloc((void **)&f->host, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            f->host = nullptr;
            f->host_bytes = 0;
            return -1;
        }
        f->host_bytes = want;
        c->fs_alloc_calls++;
        c->fs_growths++;
        if (c->fs_batches) c->fs_steady_alloc_calls++;
    }
    return 0;
}

/* Class of a candidate's interpreter storage. */
static bool prog_is_big(const FtProg &p)
{
    return p.max_stack + 1 > SMALL_S || p.max_adj + 2 > SMALL_AS || p.tape_l + 1 > SMALL_TAPE ||
           p.tape_r + 1 > SMALL_TAPE;
}

/* The row sets as the fit kernels see them: one per resident set, then the
   resident rows with the label signs in the target slot (the logistic arm's
   set, index nsets; its target is null without rc_set_fit_sign). */
std::vector<FtRows> rc_fit_rows(RcCtx *c)
{
    std::vector<FtRows> rowsd((size_t)c->nsets + 1);
    for (int s = 0; s < c->nsets; ++s) {
        rowsd[s].cols = c->sets[s].cols;
        rowsd[s].cols64 = c->sets[s].cols64;
        rowsd[s].y64 = c->sets[s].y64;
        rowsd[s].w64 = c->sets[s].w64;
        rowsd[s].n = c->sets[s].n;
        rowsd[s].v = c->sets[s].v;
    }
    rowsd[c->nsets] = rowsd[0];
    rowsd[c->nsets].y64 = c->d_tsign64;
    return rowsd;
}

/* One candidate's fit-VM streams (word pairs, trimmed to length). */
struct FtVmCode {
    int vorc = -1, vmrc = -1;
    bool vm4 = false;
    std::vector<uint32_t> vf, vr, vt;
};

/* Compile a candidate for the fit VM: the value-pass stream (value-only
   form, inputs in U, 8-slot file) and the forward/reverse pair, at 4 rows
   per lane when the program fits that block's files (an 8-slot B file, one
   saved adjoint, and a value-pass stream for the f64 polish), else at 2 rows
   per lane, whose block has the 2-row column stride.  A stream that does not
   compile leaves its code nonzero: the candidate then runs on the general
   interpreter. */
static void vm_compile(FtVmCode &o, const uint32_t *rec, int words, int p, int nv, int tune_rows,
                       bool all64)
{
    uint32_t vf[2 * 4 * FT_MAX_NODES], vr[2 * 8 * FT_MAX_NODES], vt[2 * 4 * FT_MAX_NODES];
    int info[6] = {0, 0, 0, 0, 0, 0}, tinfo[6] = {0, 0, 0, 0, 0, 0};
    const int nf = (int)(sizeof(vf) / 4), nr = (int)(sizeof(vr) / 4), ntv = (int)(sizeof(vt) / 4);
    o.vorc = rch_fitvm_compile(rec, words, p, nv, VM_XS(4) * 4,
                               VMC_INPUTS_IN_U | VMC_VALUE_ONLY | VMC_SMALL_U, vt, ntv, nullptr, 0,
                               tinfo);
    o.vmrc = rch_fitvm_compile(rec, words, p, nv, VM_XS(4) * 4, VMC_INPUTS_IN_U, vf, nf, vr, nr,
                               info);
    o.vm4 = o.vmrc == 0 && info[3] <= 8 && info[4] <= 1 && o.vorc == 0 && tune_rows != 2 && !all64;
    /* The 2-row block: f32 columns, or the plain FP64 arm's f64 columns
       (same register files, wider column stride). */
    if (o.vmrc == 0 && !o.vm4)
        o.vmrc = rch_fitvm_compile(rec, words, p, nv, all64 ? VM64_XS(2) * 8 : VM_XS(2) * 4,
                                   VMC_INPUTS_IN_U, vf, nf, vr, nr, info);
    if (o.vorc == 0) o.vt.assign(vt, vt + 2 * tinfo[0]);
    if (o.vmrc == 0) {
        o.vf.assign(vf, vf + 2 * info[0]);
        o.vr.assign(vr, vr + 2 * info[1]);
    }
}

/* The shared driver of rc_fit_batch / rc_fit_batch_hetero. */
static float fit_run(RcCtx *c, const uint32_t *prog, const uint32_t *off, const int *pcount,
                     int nprog, int kmax, const double *c0, const int *rowset_in,
                     const int *metric_in, int metric_all, int iters, double *outc,
                     double *outsse, int *outstatus)
{
    if (!c->fit) c->fit = new RcFitState();
    RcFitState *f = c->fit;
    const double t_begin = rc_now_ms();
    /* Host translation. */
    std::vector<FtProg> progs((size_t)nprog);
    std::vector<int> rowset((size_t)nprog), metric((size_t)nprog);
    bool any_irls = false;
    for (int i = 0; i < nprog; ++i) {
        const int rs = rowset_in ? rowset_in[i] : 0;
        const int mt = metric_in ? metric_in[i] : metric_all;
        if (rs < 0 || rs >= c->nsets) return -1.0f;
        if (mt != RC_LOSS_SQUARED && mt != RC_LOSS_LOG_SQUASHED && mt != RC_LOSS_ABSOLUTE &&
            mt != RC_LOSS_LOGISTIC)
            return -1.0f;
        /* The logistic arm needs the label signs of the resident rows
           (rc_set_fit_sign), like the classification scores. */
        if (mt == RC_LOSS_LOGISTIC && (rs != 0 || !c->has_tsign || !c->d_tsign64)) return -1.0f;
        rowset[i] = rs;
        metric[i] = mt;
        if (mt == RC_LOSS_ABSOLUTE) any_irls = true;
    }
    /* Per-candidate translation and fit-VM compilation are independent:
       they run on host threads, and the streams are laid out in candidate
       order afterwards. */
    /* Mode 1: the plain FP64 arm.  The same items, launch sequence and
       control kernel run every row pass on f64 rows with f64 constants: the
       assembly on the f64 kernels (FP64 MFMA Gram), the LM trial losses on
       the exact-loss kernels. */
    const bool all64 = f->mode == 1;
    std::vector<FtVmCode> vmc((size_t)nprog);
    rc_parallel_for(nprog, 32, [&](int lo, int hi) {
        for (int i = lo; i < hi; ++i) {
            const int words = (int)(off[i + 1] - off[i]);
            if (pcount[i] < 1 || pcount[i] > kmax) {
                progs[i].pre_status = 1;
                continue;
            }
            const int nv = c->sets[rowset[i]].v;
            progs[i] = translate_fit(prog + off[i], words, pcount[i], nv, true, false);
            if (progs[i].pre_status < 0)
                vm_compile(vmc[i], prog + off[i], words, pcount[i], nv, c->tune_fit_rows, all64);
        }
    });
    for (int s = 0; s < c->nsets; ++s)
        if (rc_rows_ensure_f64(c, &c->sets[s]) != 0) return -1.0f;

    /* Candidate descriptors, streams and work items. */
    std::vector<FtCand> cands((size_t)nprog);
    std::vector<uint32_t> code;
    std::vector<FtItem> items;
    std::vector<int> pre((size_t)nprog, -1);
    /* Assembly classes: fit VM with 4 rows per lane (KC16, KC32), fit VM
       with 2 rows per lane (KC16, KC32), general interpreter with the
       register stores (KC16, KC32) and with the oversized stores (KC16,
       KC32).  Value-pass classes: fit VM, general small, general big. */
    enum { NA = 8, NT = 3 };
    std::vector<int> list[NA];
    std::vector<int> tlist[NT];
    /* The trial lists split per candidate into the first (nitems + 1) / 2
       items and the rest, for the LM trial passes' early rejection. */
    std::vector<int> tlist_a[NT], tlist_b[NT];
    int maxK = 2;
    const int tile_rows = c->tune_fit_tile > 0 ? c->tune_fit_tile : TILE_ROWS;
    size_t p64_blocks = 0; /* f64 prediction buffer, in 64-row blocks */
    for (int i = 0; i < nprog; ++i) {
        FtCand &d = cands[i];
        memset(&d, 0, sizeof(d));
        d.rowset = metric[i] == RC_LOSS_LOGISTIC ? c->nsets : rowset[i];
        d.metric = metric[i];
        pre[i] = progs[i].pre_status;
        /* A host-decided candidate keeps k = 0: its outputs are c0 verbatim. */
        d.k = (pre[i] >= 0) ? 0 : pcount[i];
        if (pre[i] >= 0) continue;
        const FtVmCode &vc = vmc[i];
        const bool vm4 = vc.vm4;
        if (vc.vmrc == 0) {
            if (code.size() & 1) code.push_back(0);
            d.vm = 1;
            d.fwd_off = (uint32_t)code.size();
            code.insert(code.end(), vc.vf.begin(), vc.vf.end());
            d.rev_off = (uint32_t)code.size();
            code.insert(code.end(), vc.vr.begin(), vc.vr.end());
            d.m = (int)vc.vf.size() / 2;
            d.nrev = (int)vc.vr.size() / 2;
        } else {
            d.fwd_off = (uint32_t)code.size();
            code.insert(code.end(), progs[i].fwd.begin(), progs[i].fwd.end());
            d.rev_off = (uint32_t)code.size();
            code.insert(code.end(), progs[i].rev.begin(), progs[i].rev.end());
            d.m = (int)progs[i].fwd.size();
            d.nrev = (int)progs[i].rev.size();
        }
        int tcls;
        if (vc.vorc == 0) {
            if (code.size() & 1) code.push_back(0);
            d.t_off = (uint32_t)code.size();
            code.insert(code.end(), vc.vt.begin(), vc.vt.end());
            d.t_m = (int)vc.vt.size() / 2;
            tcls = 0;
        } else {
            d.t_off = (uint32_t)code.size();
            code.insert(code.end(), progs[i].fwd.begin(), progs[i].fwd.end());
            d.t_m = (int)progs[i].fwd.size();
            tcls = progs[i].max_stack > SMALL_S ? 2 : 1;
        }
        const int n = c->sets[rowset[i]].n;
        d.item0 = (int)items.size();
        const bool big = prog_is_big(progs[i]);
        const int kc = d.k + 2 <= 16 ? 0 : 1;
        const int cls = d.vm ? (vm4 ? 0 : 2) + kc : (big ? 6 : 4) + kc;
        maxK = std::max(maxK, d.k + 2);
        const int ntiles = (n + tile_rows - 1) / tile_rows;
        for (int r0 = 0, tix = 0; r0 < n; r0 += tile_rows, ++tix) {
            FtItem it;
            it.cand = i;
            it.row0 = r0;
            it.rows = std::min(tile_rows, n - r0);
            it.p64_blk = (int)std::min<size_t>(p64_blocks, (size_t)INT_MAX);
            p64_blocks += (size_t)(it.rows + 63) / 64;
            list[cls].push_back((int)items.size());
            tlist[tcls].push_back((int)items.size());
            (tix < (ntiles + 1) / 2 ? tlist_a : tlist_b)[tcls].push_back((int)items.size());
            items.push_back(it);
        }
        d.nitems = (int)items.size() - d.item0;
    }
    if (code.empty()) code.push_back(0);
    const int nitems = (int)items.size();
    const int tri_stride = maxK * (maxK + 1) / 2;
    /* Longest items first within each class list: fewer stragglers.  The
       sort runs on (work, item) pairs so the keys are read in place. */
    std::vector<std::pair<long long, int>> keyed;
    auto by_work = [&](std::vector<int> &l) {
        keyed.resize(l.size());
        for (size_t j = 0; j < l.size(); ++j) {
            const FtItem &it = items[(size_t)l[j]];
            const FtCand &d = cands[(size_t)it.cand];
            keyed[j] = {(long long)(d.m + d.nrev) * it.rows, l[j]};
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const std::pair<long long, int> &a, const std::pair<long long, int> &b) {
                             return a.first > b.first;
                         });
        for (size_t j = 0; j < l.size(); ++j) l[j] = keyed[j].second;
    };
    for (int q = 0; q < NA; ++q) by_work(list[q]);
    for (int q = 0; q < NT; ++q) {
        by_work(tlist[q]);
        by_work(tlist_a[q]);
        by_work(tlist_b[q]);
    }
    const std::vector<FtRows> rowsd = rc_fit_rows(c);
    std::vector<int> flat_lists;
    int lofs[NA + 1], tofs[NT + 1];
    for (int q = 0; q < NA; ++q) {
        lofs[q] = (int)flat_lists.size();
        flat_lists.insert(flat_lists.end(), list[q].begin(), list[q].end());
    }
    lofs[NA] = (int)flat_lists.size();
    for (int q = 0; q < NT; ++q) {
        tofs[q] = (int)flat_lists.size();
        flat_lists.insert(flat_lists.end(), tlist[q].begin(), tlist[q].end());
    }
    tofs[NT] = (int)flat_lists.size();
    int tofs_a[NT + 1], tofs_b[NT + 1];
    for (int q = 0; q < NT; ++q) {
        tofs_a[q] = (int)flat_lists.size();
        flat_lists.insert(flat_lists.end(), tlist_a[q].begin(), tlist_a[q].end());
    }
    tofs_a[NT] = (int)flat_lists.size();
    for (int q = 0; q < NT; ++q) {
        tofs_b[q] = (int)flat_lists.size();
        flat_lists.insert(flat_lists.end(), tlist_b[q].begin(), tlist_b[q].end());
    }
    tofs_b[NT] = (int)flat_lists.size();
    if (flat_lists.empty()) flat_lists.push_back(0);

    /* Region layout: inputs [code | cands | rows | items | lists | pre | c0]
       then outputs [c | sse | status] then work. */
    /* The exact passes of the polish store every row's f64 prediction for
       the polish assembly that follows at the same constants (4-row fit VM
       classes), while the buffer stays within FT_P64_CAP; beyond it the
       polish assembly computes the prediction itself. */
    size_t p64_cap = FT_P64_CAP;
    if (const char *e = getenv("RCSR_FIT_P64_CAP")) p64_cap = (size_t)strtoull(e, nullptr, 10);
    const bool use_p64 = !all64 && p64_blocks <= (size_t)INT_MAX &&
                         p64_blocks * 64 * sizeof(double) <= p64_cap;
    const FtRegion rg = ft_region(code.size(), cands.size(), rowsd.size(), items.size(),
                                  flat_lists.size(), nprog, kmax, tri_stride,
                                  use_p64 ? p64_blocks * 64 * sizeof(double) : 0);
    const size_t o_code = rg.o_code, o_cand = rg.o_cand, o_rows = rg.o_rows, o_items = rg.o_items;
    const size_t o_lists = rg.o_lists, o_pre = rg.o_pre, o_c0 = rg.o_c0, in_end = rg.in_end;
    const size_t o_outc = rg.o_outc, o_sse = rg.o_sse, o_st = rg.o_st, io_end = rg.io_end;
    const size_t o_state = rg.o_state, o_run_a = rg.o_run_a, o_run_t = rg.o_run_t;
    const size_t o_cp32 = rg.o_cp32, o_cp64 = rg.o_cp64, o_pgram = rg.o_pgram, o_ploss = rg.o_ploss;
    const size_t o_pbad = rg.o_pbad, o_pskip = rg.o_pskip, dev_end = rg.dev_end;
    double t = rc_now_ms();
    if (fit_reserve(c, dev_end, io_end) != 0) return -1.0f;
    c->stats.alloc_ms += rc_now_ms() - t;

    unsigned char *H = f->host;
    unsigned char *D = f->dev;
    memcpy(H + o_code, code.data(), code.size() * 4);
    memcpy(H + o_cand, cands.data(), cands.size() * sizeof(FtCand));
    memcpy(H + o_rows, rowsd.data(), rowsd.size() * sizeof(FtRows));
    if (!items.empty()) memcpy(H + o_items, items.data(), items.size() * sizeof(FtItem));
    memcpy(H + o_lists, flat_lists.data(), flat_lists.size() * 4);
    memcpy(H + o_pre, pre.data(), (size_t)nprog * 4);
    memcpy(H + o_c0, c0, (size_t)nprog * kmax * 8);
    t = rc_now_ms();
    hipStream_t s = c->s_fit;
    if (hipMemcpyAsync(D, H, in_end, hipMemcpyHostToDevice, s) != hipSuccess) return -1.0f;
    c->stats.h2d_ms += rc_now_ms() - t;
    c->stats.h2d_bytes += in_end;

    const uint32_t *d_code = (const uint32_t *)(D + o_code);
    const FtCand *d_cand = (const FtCand *)(D + o_cand);
    const FtRows *d_rows = (const FtRows *)(D + o_rows);
    const FtItem *d_items = (const FtItem *)(D + o_items);
    const int *d_lists = (const int *)(D + o_lists);
    const double *d_c0 = (const double *)(D + o_c0);
    double *d_outc = (double *)(D + o_outc);
    double *d_sse = (double *)(D + o_sse);
    int *d_st = (int *)(D + o_st);
    FtState *d_state = (FtState *)(D + o_state);
    int *d_run_a = (int *)(D + o_run_a);
    int *d_run_t = (int *)(D + o_run_t);
    int *d_run_b = (int *)(D + rg.o_run_b);
    float *d_cp32 = (float *)(D + o_cp32);
    double *d_cp64 = (double *)(D + o_cp64);
    double *d_p64 = use_p64 ? (double *)(D + rg.o_p64) : nullptr;
    double *d_pgram = (double *)(D + o_pgram);
    double *d_ploss = (double *)(D + o_ploss);
    int *d_pbad = (int *)(D + o_pbad);
    int *d_pskip = (int *)(D + o_pskip);

    if ((!f->side && hipStreamCreateWithFlags(&f->side, hipStreamNonBlocking) != hipSuccess) ||
        (!f->ev_fork && hipEventCreateWithFlags(&f->ev_fork, hipEventDisableTiming) != hipSuccess) ||
        (!f->ev_join && hipEventCreateWithFlags(&f->ev_join, hipEventDisableTiming) != hipSuccess)) {
        (void)hipGetLastError();
        return -1.0f;
    }

    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> marks;
    auto mark = [&](int kind, auto &&fn) {
        if (!f->profile) { fn(); return; }
        hipEvent_t a, b;
        (void)hipEventCreate(&a);
        (void)hipEventCreate(&b);
        (void)hipEventRecord(a, s);
        fn();
        (void)hipEventRecord(b, s);
        marks.push_back({kind, {a, b}});
    };
    auto ctrl0 = [&](int op) {
        ft_ctrl_kernel<<<nprog, 64, 0, s>>>(op, nprog, d_state, d_cand, d_c0, kmax, d_run_a, d_run_t,
                                            d_cp32, d_cp64, d_pgram, tri_stride, d_ploss, d_pbad,
                                            d_pskip, d_outc, d_sse, d_st, all64 ? 1 : 0);
    };
    auto ctrl = [&](int op) { mark(3, [&]() { ctrl0(op); }); };
    /* The classes of a pass write disjoint items: the large ones run on
       the fit stream, the small ones on the side stream between a fork and
       a join, so the small classes fill the slots the large ones leave. */
    hipStream_t &sd = f->side;
    auto fork = [&]() {
        (void)hipEventRecord(f->ev_fork, s);
        (void)hipStreamWaitEvent(sd, f->ev_fork, 0);
    };
    auto join = [&]() {
        (void)hipEventRecord(f->ev_join, sd);
        (void)hipStreamWaitEvent(s, f->ev_join, 0);
    };
    auto asm_launch = [&](auto kern, int q, hipStream_t st) {
        const int nl = lofs[q + 1] - lofs[q];
        if (nl <= 0) return;
        kern<<<nl, 64, 0, st>>>(d_items, d_lists + lofs[q], nl, d_cand, d_rows, d_code, d_run_a,
                                d_cp32, d_cp64, d_pgram, tri_stride, d_ploss, d_pbad, d_pskip);
    };
    auto vm_asm = [&](int kc, int r, bool polish, int q, hipStream_t st) {
        vm_launch_asm(kc, r, polish, lofs[q + 1] - lofs[q], st, d_items, d_lists + lofs[q], d_cand,
                      d_rows, d_code, d_run_a, d_cp32, d_cp64, d_pgram, tri_stride, d_ploss,
                      d_pbad, d_pskip, d_p64);
    };
    auto asm64_launch = [&](auto kern, int q, hipStream_t st) {
        const int nl = lofs[q + 1] - lofs[q];
        if (nl <= 0) return;
        kern<<<nl, 64, 0, st>>>(d_items, d_lists + lofs[q], nl, d_cand, d_rows, d_code, d_run_a,
                                d_cp64, d_pgram, tri_stride, d_ploss, d_pbad, d_pskip);
    };
    auto vm_asm64 = [&](int kc, bool polish, int q, hipStream_t st) {
        vm_launch_asm64(kc, polish, lofs[q + 1] - lofs[q], st, d_items, d_lists + lofs[q], d_cand,
                        d_rows, d_code, d_run_a, d_cp64, d_pgram, tri_stride, d_ploss, d_pbad,
                        d_pskip);
    };
    auto asm_pass0 = [&](int polish) {
        fork();
        if (all64) {
            /* Classes 0 and 1 (4 f32 rows per lane) are empty here. */
            vm_asm64(16, polish != 0, 2, s);
            vm_asm64(32, polish != 0, 3, s);
            if (polish) {
                asm64_launch(ft_asm64_kernel<16, false, true>, 4, sd);
                asm64_launch(ft_asm64_kernel<32, false, true>, 5, sd);
                asm64_launch(ft_asm64_kernel<16, true, true>, 6, sd);
                asm64_launch(ft_asm64_kernel<32, true, true>, 7, sd);
            } else {
                asm64_launch(ft_asm64_kernel<16, false, false>, 4, sd);
                asm64_launch(ft_asm64_kernel<32, false, false>, 5, sd);
                asm64_launch(ft_asm64_kernel<16, true, false>, 6, sd);
                asm64_launch(ft_asm64_kernel<32, true, false>, 7, sd);
            }
        } else if (polish) {
            vm_asm(16, 4, true, 0, s);
            vm_asm(32, 4, true, 1, s);
            vm_asm(16, 2, true, 2, sd);
            vm_asm(32, 2, true, 3, sd);
            asm_launch(ft_asm_kernel<16, false, true>, 4, sd);
            asm_launch(ft_asm_kernel<32, false, true>, 5, sd);
            asm_launch(ft_asm_kernel<16, true, true>, 6, sd);
            asm_launch(ft_asm_kernel<32, true, true>, 7, sd);
        } else {
            vm_asm(16, 4, false, 0, s);
            vm_asm(32, 4, false, 1, s);
            vm_asm(16, 2, false, 2, sd);
            vm_asm(32, 2, false, 3, sd);
            asm_launch(ft_asm_kernel<16, false, false>, 4, sd);
            asm_launch(ft_asm_kernel<32, false, false>, 5, sd);
            asm_launch(ft_asm_kernel<16, true, false>, 6, sd);
            asm_launch(ft_asm_kernel<32, true, false>, 7, sd);
        }
        join();
    };
    auto asm_pass = [&](int polish) { mark(0, [&]() { asm_pass0(polish); }); };
    /* One trial pass over the item lists at offsets of (full, first
       halves or second halves) with the run flags run. */
    auto trial_lists = [&](bool exact_, const int *of, const int *run) {
        const bool exact = exact_ || all64;
        fork();
        for (int q = 0; q < NT; ++q) {
            const int nl = of[q + 1] - of[q];
            if (nl <= 0) continue;
            const int *L = d_lists + of[q];
            const hipStream_t st = q == 0 ? s : sd;
            if (!exact) {
                if (q == 0)
                    vm_launch_trial(false, nl, st, d_items, L, d_cand, d_rows, d_code, run, d_cp32, d_cp64, d_ploss,
                                    nullptr);
                else if (q == 1)
                    ft_trial_kernel<float, false><<<nl, 64, 0, st>>>(d_items, L, nl, d_cand, d_rows, d_code, run, d_cp32, d_ploss);
                else
                    ft_trial_kernel<float, true><<<nl, 64, 0, st>>>(d_items, L, nl, d_cand, d_rows, d_code, run, d_cp32, d_ploss);
            } else {
                if (q == 0)
                    vm_launch_trial(true, nl, st, d_items, L, d_cand, d_rows, d_code, run, d_cp32, d_cp64, d_ploss,
                                    d_p64);
                else if (q == 1)
                    ft_trial_kernel<double, false><<<nl, 64, 0, st>>>(d_items, L, nl, d_cand, d_rows, d_code, run, d_cp64, d_ploss);
                else
                    ft_trial_kernel<double, true><<<nl, 64, 0, st>>>(d_items, L, nl, d_cand, d_rows, d_code, run, d_cp64, d_ploss);
            }
        }
        join();
    };

    auto trial_pass0 = [&](bool exact) { trial_lists(exact, tofs, d_run_t); };
    /* An LM trial pass: the first halves, the prune, the second halves of
       the candidates not yet rejected. */
    auto trial_lm0 = [&]() {
        trial_lists(false, tofs_a, d_run_t);
        ft_trial_prune<<<(nprog + 63) / 64, 64, 0, s>>>(nprog, d_state, d_cand, d_ploss, d_run_t,
                                                         d_run_b);
        trial_lists(false, tofs_b, d_run_b);
    };
    auto trial_pass = [&](bool exact) { mark(exact ? 2 : 1, [&]() { trial_pass0(exact); }); };
    auto trial_lm = [&]() { mark(1, [&]() { trial_lm0(); }); };

    /* The launch sequence: recorded as a hipGraph per batch shape. */
    std::vector<uint64_t> key = ft_graph_key(nprog, nitems, iters, any_irls, kmax, tri_stride, D,
                                             rg, lofs, NA + 1, tofs, NT + 1, f->mode);
    for (int q = 0; q <= NT; ++q) {
        key.push_back((uint64_t)tofs_a[q]);
        key.push_back((uint64_t)tofs_b[q]);
    }
    const char *nograph = getenv("RCSR_FIT_NO_GRAPH");
    const bool use_graph = !(nograph && nograph[0] == '1') && !f->profile;
    auto record = [&]() {
        ft_state_init<<<(nprog + 63) / 64, 64, 0, s>>>(d_state, nprog, (const int *)(D + o_pre), iters);
        ctrl(CT_INIT);
        for (int slot = 0; slot < iters; ++slot) {
            asm_pass(0);
            ctrl(CT_AFTER_ASM);
            trial_lm();
            ctrl(CT_AFTER_TRIAL);
            if (any_irls) {
                for (int h = 0; h < 2; ++h) {
                    trial_lm();
                    ctrl(CT_AFTER_TRIAL);
                }
            }
        }
        ctrl(CT_PRE_REBASE);
        trial_pass(true);
        ctrl(CT_REBASE);
        for (int pit = 0; pit < 4; ++pit) {
            asm_pass(1);
            ctrl(CT_AFTER_ASM_P);
            trial_pass(true);
            ctrl(CT_AFTER_TRIAL_P);
            if (any_irls) {
                trial_pass(true);
                ctrl(CT_AFTER_TRIAL_P);
            }
        }
        ctrl(CT_FINAL);
    };
    t = rc_now_ms();
    (void)hipEventRecord(c->t0, s);
    if (use_graph) {
        if (!f->exec || f->graph_key != key) {
            if (f->exec) { (void)hipGraphExecDestroy(f->exec); f->exec = nullptr; }
            if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
            hipStream_t cap;
            if (hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) != hipSuccess) return -1.0f;
            std::swap(s, cap);
            bool okc = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
            if (okc) {
                record();
                okc = hipStreamEndCapture(s, &f->graph) == hipSuccess;
            }
            std::swap(s, cap);
            (void)hipStreamDestroy(cap);
            if (!okc || hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGetLastError();
                if (f->graph) { (void)hipGraphDestroy(f->graph); f->graph = nullptr; }
                f->exec = nullptr;
                return -1.0f;
            }
            f->graph_key = key;
        }
        if (hipGraphLaunch(f->exec, s) != hipSuccess) return -1.0f;
    } else {
        record();
    }
    (void)hipEventRecord(c->t1, s);
    c->stats.launch_ms += rc_now_ms() - t;
    if (hipGetLastError() != hipSuccess) return -1.0f;
    t = rc_now_ms();
    if (hipMemcpyAsync(H + in_end, D + in_end, io_end - in_end, hipMemcpyDeviceToHost, s) != hipSuccess)
        return -1.0f;
    if (hipStreamSynchronize(s) != hipSuccess || hipGetLastError() != hipSuccess) return -1.0f;
    c->stats.sync_ms += rc_now_ms() - t;
    for (auto &mk : marks) {
        float e = 0.0f;
        (void)hipEventElapsedTime(&e, mk.second.first, mk.second.second);
        f->t_pass[mk.first] += e;
        f->n_pass[mk.first] += 1;
        (void)hipEventDestroy(mk.second.first);
        (void)hipEventDestroy(mk.second.second);
    }
    f->last_state = d_state;
    f->last_nprog = nprog;
    memcpy(outc, H + o_outc, (size_t)nprog * kmax * 8);
    memcpy(outsse, H + o_sse, (size_t)nprog * 8);
    memcpy(outstatus, H + o_st, (size_t)nprog * 4);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    c->stats.kernel_ms = ms;
    c->stats.batch = nprog;
    c->stats.wall_ms = rc_now_ms() - t_begin;
    c->fs_batches++;
    return ms;
}

extern "C" int rch_fit_graph_key(int code_words, int ncands, int nrowsets, int nitems,
                                 int nlist_words, int nprog, int kmax, int tri_stride,
                                 unsigned long long *key, int cap)
{
    if (code_words < 0 || ncands < 0 || nrowsets < 0 || nitems < 0 || nlist_words < 0 ||
        nprog <= 0 || kmax < 0 || tri_stride < 0 || !key)
        return -1;
    const FtRegion g = ft_region((size_t)code_words, (size_t)ncands, (size_t)nrowsets,
                                 (size_t)nitems, (size_t)nlist_words, nprog, kmax, tri_stride);
    const int zeros[1] = {0};
    const std::vector<uint64_t> k =
        ft_graph_key(nprog, nitems, 30, 0, kmax, tri_stride, nullptr, g, zeros, 1, zeros, 1, 0);
    if ((int)k.size() > cap) return -2;
    for (size_t q = 0; q < k.size(); ++q) key[q] = (unsigned long long)k[q];
    return (int)k.size();
}

extern "C" {

float rc_fit_batch(void *ctx, const unsigned int *prog, const unsigned int *off, const int *pcount,
                   int nprog, int kmax, const double *c0, const int *gauge, int iters, int metric,
                   double *outc, double *outsse, int *outstatus)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || !off || !pcount || !c0 || nprog <= 0 || kmax <= 0 || iters < 0) return -1.0f;
    if (!outc || !outsse || !outstatus) return -1.0f;
    if (kmax > FT_KMAX) return -1.0f;
    if (metric == RC_LOSS_CORRELATION || metric == RC_LOSS_CORRELATION_SIGNED) {
        /* The correlation arm with the gauge pairs (eval.cu:2320-2448). */
        if (rc_corr_run(c, prog, off, pcount, nprog, kmax, c0, gauge, iters, metric, outc, outsse,
                        outstatus) != 0)
            return -1.0f;
        return (float)c->stats.kernel_ms;
    }
    return fit_run(c, prog, off, pcount, nprog, kmax, c0, nullptr, nullptr, metric, iters, outc,
                   outsse, outstatus);
}

float rc_fit_batch_hetero(void *ctx, const unsigned int *prog, const unsigned int *off,
                          const int *pcount, int nprog, int kmax, const double *c0,
                          const int *rowset, const int *metric, int iters, double *outc,
                          double *outsse, int *outstatus)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || !off || !pcount || !c0 || !rowset || !metric || nprog <= 0 || kmax <= 0 ||
        iters < 0 || !outc || !outsse || !outstatus)
        return -1.0f;
    if (kmax > FT_KMAX) return -1.0f;
    return fit_run(c, prog, off, pcount, nprog, kmax, c0, rowset, metric, 0, iters, outc, outsse,
                   outstatus);
}

int rc_fit_kernel_info(int kmax, int *threads, int *smem_bytes, int *max_constants)
{
    if (!threads || !smem_bytes || !max_constants) return -1;
    if (kmax <= 0 || kmax > FT_KMAX) return -1;
    *threads = 64;
    *smem_bytes = (kmax + 2 <= 16 ? 16 : 32) * XS * 4;
    *max_constants = FT_KMAX;
    return 0;
}

int rc_fit_profile(void *ctx, int enable)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return -1;
    if (!c->fit) c->fit = new RcFitState();
    const int prev = c->fit->profile;
    c->fit->profile = enable ? 1 : 0;
    for (int q = 0; q < 4; ++q) {
        c->fit->t_pass[q] = 0.0;
        c->fit->n_pass[q] = 0;
    }
    return prev;
}

int rc_fit_report(void *ctx, RcFitReport *out, int *per_cand, int nprog)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || !c->fit) return -1;
    RcFitState *f = c->fit;
    out->asm_ms = f->t_pass[0];
    out->trial_ms = f->t_pass[1];
    out->exact_ms = f->t_pass[2];
    out->ctrl_ms = f->t_pass[3];
    out->asm_launches = f->n_pass[0];
    out->trial_launches = f->n_pass[1];
    out->exact_launches = f->n_pass[2];
    out->ctrl_launches = f->n_pass[3];
    if (per_cand) {
        if (!f->last_state || nprog != f->last_nprog) return -1;
        std::vector<FtState> h((size_t)nprog);
        if (hipMemcpy(h.data(), f->last_state, sizeof(FtState) * (size_t)nprog,
                      hipMemcpyDeviceToHost) != hipSuccess)
            return -1;
        for (int i = 0; i < nprog; ++i) {
            per_cand[5 * i + 0] = h[i].n_asm;
            per_cand[5 * i + 1] = h[i].n_trial;
            per_cand[5 * i + 2] = h[i].n_exact;
            per_cand[5 * i + 3] = h[i].acc_bits;
            per_cand[5 * i + 4] = h[i].n_decided;
        }
    }
    return 0;
}

int rc_set_fit_mode(void *ctx, int mode)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || (mode != 0 && mode != 1)) return -1;
    if (!c->fit) c->fit = new RcFitState();
    const int prev = c->fit->mode;
    c->fit->mode = mode;
    return prev;
}

int rc_fit_jacobian(void *ctx, const unsigned int *prog, const unsigned int *off, const int *pcount,
                    int nprog, int kmax, const double *cvals, int cand, int rows, float *out_pred,
                    float *out_jac, float *out_gx)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !prog || !off || !pcount || !cvals || cand < 0 || cand >= nprog || rows <= 0 ||
        kmax <= 0 || kmax > FT_KMAX || !out_pred || !out_jac || !out_gx)
        return -1;
    RcRows *R = &c->sets[0];
    if (R->v > 8) return -1;
    rows = std::min(rows, R->n);
    FtProg p = translate_fit(prog + off[cand], (int)(off[cand + 1] - off[cand]), pcount[cand], R->v,
                             true, true);
    if (p.pre_status >= 0) return -1;
    if (p.max_stack + 1 > BIG_S || p.max_adj + 2 > BIG_AS || p.tape_l + 1 > BIG_TAPE ||
        p.tape_r + 1 > BIG_TAPE)
        return -1;
    std::vector<uint32_t> code = p.fwd;
    FtCand d;
    memset(&d, 0, sizeof(d));
    d.fwd_off = 0;
    d.rev_off = (uint32_t)code.size();
    code.insert(code.end(), p.rev.begin(), p.rev.end());
    d.m = (int)p.fwd.size();
    d.nrev = (int)p.rev.size();
    d.k = pcount[cand];
    FtRows rd;
    rd.cols = R->cols;
    rd.cols64 = R->cols64;
    rd.y64 = R->y64;
    rd.w64 = R->w64;
    rd.n = R->n;
    rd.v = R->v;
    std::vector<float> cp(FT_KMAX, 0.0f);
    for (int j = 0; j < d.k; ++j) cp[j] = (float)cvals[(size_t)cand * kmax + j];
    uint32_t *dc = nullptr;
    FtCand *dd = nullptr;
    FtRows *dr = nullptr;
    float *dcp = nullptr, *dp = nullptr, *dj = nullptr, *dg = nullptr;
    int rc = -1;
    if (hipMalloc((void **)&dc, code.size() * 4) == hipSuccess &&
        hipMalloc((void **)&dd, sizeof(FtCand)) == hipSuccess &&
        hipMalloc((void **)&dr, sizeof(FtRows)) == hipSuccess &&
        hipMalloc((void **)&dcp, FT_KMAX * 4) == hipSuccess &&
        hipMalloc((void **)&dp, (size_t)rows * 4) == hipSuccess &&
        hipMalloc((void **)&dj, (size_t)rows * kmax * 4) == hipSuccess &&
        hipMalloc((void **)&dg, (size_t)rows * std::max(R->v, 1) * 4) == hipSuccess) {
        (void)hipMemcpy(dc, code.data(), code.size() * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(dd, &d, sizeof(FtCand), hipMemcpyHostToDevice);
        (void)hipMemcpy(dr, &rd, sizeof(FtRows), hipMemcpyHostToDevice);
        (void)hipMemcpy(dcp, cp.data(), FT_KMAX * 4, hipMemcpyHostToDevice);
        (void)hipMemset(dj, 0, (size_t)rows * kmax * 4);
        ft_probe_kernel<<<(rows + 63) / 64, 64, 0, c->s_fit>>>(dd, dr, dc, dcp, 0, rows, dp, dj, kmax, dg);
        if (hipStreamSynchronize(c->s_fit) == hipSuccess && hipGetLastError() == hipSuccess) {
            (void)hipMemcpy(out_pred, dp, (size_t)rows * 4, hipMemcpyDeviceToHost);
            (void)hipMemcpy(out_jac, dj, (size_t)rows * kmax * 4, hipMemcpyDeviceToHost);
            (void)hipMemcpy(out_gx, dg, (size_t)rows * R->v * 4, hipMemcpyDeviceToHost);
            rc = 0;
        }
    }
    (void)hipGetLastError();
    if (dc) (void)hipFree(dc);
    if (dd) (void)hipFree(dd);
    if (dr) (void)hipFree(dr);
    if (dcp) (void)hipFree(dcp);
    if (dp) (void)hipFree(dp);
    if (dj) (void)hipFree(dj);
    if (dg) (void)hipFree(dg);
    return rc;
}

} /* extern "C" */

