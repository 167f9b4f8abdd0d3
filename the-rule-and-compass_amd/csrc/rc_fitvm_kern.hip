/*
 * The fit VM kernels: assembly (forward + adjoint + FP32-MFMA Gram) and the
 * value passes (trial and exact losses), compiled in their own translation
 * unit.  They are single-wave interpreters dispatching on wave-uniform
 * instruction words, so this file is built with uniform regions left
 * unstructured (a dispatch test is one s_bitcmp and one branch); every
 * other kernel of the library keeps the default structurization.
 *
 * Reference behaviour: ft_forward / ft_backward (eval.cu:1173-1317),
 * ft_assemble (1693-1804) and ft_loss (1809-1855); the instruction streams
 * come from the fit VM compiler (rc_fitvm.cpp).
 */
#include "rc_fit_dev.h"
#include "rc_fitvm_dev.h"
#include "rc_fitvm_kern.h"

#include <math.h>

namespace {

/* Per-row residual bookkeeping shared by the assembly kernels. */
struct RowPrep {
    double r64, s64, rho, wb;
    bool live, ok;
};

/* UNIT: the squared metric with no row weights, where the weight is 1.0
   and its square root exactly 1.0 (so the f64 root is not computed). */
template <bool UNIT>
__device__ __forceinline__ RowPrep vm_row_prep(bool live, double wb, float pred, double p64,
                                               double y64, int metric, bool polish)
{
    RowPrep q;
    q.live = live;
    q.wb = wb;
    const double d64 = p64 - y64;
    q.ok = live && isfinite(p64) && (polish || fabsf(pred) <= 1e30f) && isfinite(d64);
    if (UNIT) {
        q.r64 = d64;
        q.s64 = 1.0;
        q.rho = rc_row_rho(d64, RC_LOSS_SQUARED);
    } else {
        const RcRowFit f = rc_row_fit(wb, p64, y64, metric);
        q.r64 = f.r;
        q.s64 = f.s;
        q.rho = f.rho;
    }
    return q;
}

/* Row-pack shapes of the assembly: R rows per lane, a block of 64 R rows
   and the register files. */
template <int R> struct VmAsmShape;
template <> struct VmAsmShape<2> {
    typedef vm_f2 V;
    typedef vm_u64x16 FU;
    typedef vm_u64x16 FB;
    typedef vm_u64x8 FA;
};
template <> struct VmAsmShape<4> {
    typedef vm_f4 V;
    typedef vm_u64x16 FU;
    typedef vm_u64x8 FB;
    typedef vm_u64x2 FA;
};

/* Assembly on the fit VM: the same per-row contract as ft_asm_kernel (the
   augmented row [s*J | s*r_hi | s*r_lo] of every row into the Gram block,
   the loss sum, bad-row and skipped-row flags), 64 R rows per wave step.
   The streams hold the inputs in U slots 0..v-1.  For KC = 32 only the
   K <= 22 live columns are stored, and the Gram takes two MFMAs per four
   rows (the column pairing of ft_pair_store) where three 16 x 16 blocks
   of a 32 x 32 Gram would take three; operand reads beyond the stored
   columns are clamped to the last one and land only in discarded entries.
   The LDS block allows at most two waves per SIMD.  PM selects the polish:
   0 none; 1 the f64 prediction computed here (on the value-pass stream for
   R = 4, the tape stream for R = 2), whose sweep needs more registers than
   two waves hold, still held to two (it spills a few hundred bytes per
   lane, and measured 1.9 ms faster per C3 fit than at one wave, with the
   same bits); 2 (R = 4) the f64 prediction read from p64in, where the exact
   pass at the same constants stored what PM = 1 computes. */
template <int KC, int PM, int R>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(PM != 0 ? 2 : 1))) void vm_asm_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run, const float *__restrict__ cp32,
    const double *__restrict__ cp64, double *__restrict__ pgram, int tri_stride,
    double *__restrict__ ploss, int *__restrict__ pbad, int *__restrict__ pskip,
    const double *__restrict__ p64in)
{
    constexpr bool POLISH = PM != 0;
    static_assert(PM != 2 || R == 4, "stored predictions serve the 4-row classes");
    typedef VmAsmShape<R> S;
    typedef typename S::V V;
    typedef typename S::FU FU;
    typedef typename S::FB FB;
    typedef typename S::FA FA;
    typedef VmPack<V> P;
    constexpr int XS = VM_XS(R);
    constexpr int KCA = (KC == 32) ? FT_KMAX + 2 : KC; /* stored columns */
    static_assert(KC != 32 || KCA == 22, "the KC = 32 column pairing covers 22 columns");
    constexpr int BLOCK = 64 * R;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const FtCand d = cands[cand];
    const FtRows Rw = rowsd[d.rowset];
    const int lane = threadIdx.x;
    const int k = d.k;

    __shared__ float X[KCA * XS];
    for (int i = lane; i < KCA * XS; i += 64) X[i] = 0.0f;
    __syncthreads();

    const float *cst = cp32 + (size_t)cand * FT_KMAX;
    const double *cst64 = cp64 + (size_t)cand * FT_KMAX;
    const uint2 *fc = (const uint2 *)(code + d.fwd_off);
    const uint2 *rc = (const uint2 *)(code + d.rev_off);
    char *xl = (char *)(X + lane);
    /* KC = 16: acc00 is the Gram block.  KC = 32: acc00 and acc10 are the
       products P and Q of ft_pair_store<6>. */
    floatx4 acc00 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0};
    double g00[4] = {0, 0, 0, 0}, g10[4] = {0, 0, 0, 0};
    double lsum = 0.0;
    int bad = 0, skip = 0;
    FU U0 = (FU)(0), U1 = U0;
    FB B0 = (FB)(0), B1 = B0;
    FA A0 = (FA)(0), A1 = A0;
    const int row_end = it.row0 + it.rows;
    const bool unit_w = Rw.w64 == nullptr && d.metric == RC_LOSS_SQUARED;
    int step = 0;
    for (int base = it.row0; base < row_end; base += BLOCK, ++step) {
        int q[R];
        bool in[R];
        double wb[R], yv[R];
        /* The step's weights and targets are loaded with its inputs, so
           their latency overlaps the input loads. */
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + 64 * r + lane;
            in[r] = row < row_end;
            q[r] = in[r] ? row : it.row0;
            wb[r] = Rw.w64 ? Rw.w64[q[r]] : 1.0;
            yv[r] = Rw.y64[q[r]];
        }
#pragma unroll
        for (int v = 0; v < VM_MAX_VARS; ++v)
            if (v < Rw.v) {
                const float *col = Rw.cols + (size_t)v * (size_t)Rw.n;
                V xv;
#pragma unroll
                for (int r = 0; r < R; ++r) xv[r] = col[q[r]];
                P::set(U0, U1, U1, U1, v, xv);
            }
        const V pred = vm_forward<V, true>(fc, d.m, cst, U0, U1, U1, U1, U0, U1, U1, U1, B0, B1, B1,
                                           B1);
        double p64[R];
#pragma unroll
        for (int r = 0; r < R; ++r) p64[r] = (double)pred[r];
        if (PM == 2) {
            const double *pv = p64in + (size_t)it.p64_blk * 64 - it.row0;
#pragma unroll
            for (int r = 0; r < R; ++r) p64[r] = pv[q[r]];
        } else if (POLISH && R == 2) {
            /* f64 residual on the tape stream (16-slot file, 2 rows) */
            dv16 Ua = (dv16)(0.0), Ub = Ua, Ba = Ua, Bb = Ua;
#pragma unroll
            for (int v = 0; v < VM_MAX_VARS; ++v)
                if (v < Rw.v) {
                    const double *col = Rw.cols64 + (size_t)v * (size_t)Rw.n;
                    VmPack<vm_d2>::set(Ua, Ub, Ub, Ub, v, (vm_d2){col[q[0]], col[q[1 % R]]});
                }
            const vm_d2 pp = vm_forward<vm_d2, false>(fc, d.m, cst64, Ua, Ub, Ub, Ub, Ua, Ub, Ub,
                                                      Ub, Ba, Bb, Bb, Bb);
            p64[0] = pp.x;
            p64[1 % R] = pp.y;
        } else if (POLISH) {
            /* f64 residual on the value-pass stream (8-slot file, 4 rows) */
            const uint2 *tc = (const uint2 *)(code + d.t_off);
            dv8 W0 = (dv8)(0.0), W1 = W0, W2 = W0, W3 = W0, Wb = W0;
#pragma unroll
            for (int v = 0; v < VM_MAX_VARS; ++v)
                if (v < Rw.v) {
                    const double *col = Rw.cols64 + (size_t)v * (size_t)Rw.n;
                    VmPack<vm_d4>::set(W0, W1, W2, W3, v,
                                       (vm_d4){col[q[0]], col[q[1 % R]], col[q[2 % R]], col[q[3 % R]]});
                }
            const vm_d4 pp = vm_forward<vm_d4, false>(tc, d.t_m, cst64, W0, W1, W2, W3, W0, W1, W2,
                                                      W3, Wb, Wb, Wb, Wb);
#pragma unroll
            for (int r = 0; r < R; ++r) p64[r] = pp[r % 4];
        }
        RowPrep a[R];
        V seed;
        if (unit_w) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                a[r] = vm_row_prep<true>(in[r] && wb[r] > 0.0, wb[r], pred[r], p64[r], yv[r],
                                         d.metric, POLISH);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                a[r] = vm_row_prep<false>(in[r] && wb[r] > 0.0, wb[r], pred[r], p64[r], yv[r],
                                          d.metric, POLISH);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) seed[r] = a[r].ok ? (float)a[r].s64 : 0.0f;
        const V fin = vm_reverse<V, FU, FB, FA>(rc, d.nrev, seed, cst, U0, U1, B0, B1, A0, A1, xl);
        bool fix = false;
        bool use[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool gf = fin[r] == 0.0f;
            use[r] = a[r].ok && gf;
            if (POLISH) {
                if (a[r].live && !a[r].ok) bad = 1;
                skip += (a[r].ok && !gf) ? 1 : 0;
                if (a[r].ok) lsum += a[r].wb * a[r].rho;
            } else {
                if (a[r].live && !use[r]) bad = 1;
                if (use[r]) lsum += a[r].wb * a[r].rho;
            }
            fix = fix || (a[r].ok && !gf);
        }
        if (__any(fix)) {
            for (int j = 0; j < k; ++j)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (!use[r]) X[j * XS + 64 * r + lane] = 0.0f;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double rs = a[r].s64 * a[r].r64;
            const float rh = (float)rs;
            X[k * XS + 64 * r + lane] = use[r] ? rh : 0.0f;
            X[(k + 1) * XS + 64 * r + lane] = use[r] ? (float)(rs - (double)rh) : 0.0f;
        }
        __syncthreads();
        const int col = lane & 15;
        const int sub = lane >> 4;
        const int colc = col < 8 ? col : (8 + col < KCA ? 8 + col : KCA - 1);
        /* Double-buffered operands: the LDS reads of the next GQ groups are
           issued before the current GQ run (the scheduling fences keep them
           there), so no MFMA waits on a fresh LDS read.  The MFMAs see the
           same operands in the same order. */
        constexpr int NQ = BLOCK / 4, GQ = KC == 32 ? 4 : 8;
        static_assert(NQ % (2 * GQ) == 0, "operand batch pairs tile the block");
        const float *x0 = X + col * XS + sub;
        const float *x1 = X + (6 + col) * XS + sub;
        const float *x2 = X + colc * XS + sub;
        float p0[GQ], p1[GQ], p2[GQ], q0[GQ], q1[GQ], q2[GQ];
        auto load = [&](float *d0, float *d1, float *d2, int g) {
#pragma unroll
            for (int i = 0; i < GQ; ++i) {
                d0[i] = x0[4 * (g + i)];
                d1[i] = KC == 32 ? x1[4 * (g + i)] : 0.0f;
                d2[i] = KC == 32 ? x2[4 * (g + i)] : 0.0f;
            }
        };
        auto gram = [&](const float *b0, const float *b1, const float *b2) {
#pragma unroll
            for (int i = 0; i < GQ; ++i) {
                if (KC == 32) {
                    acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[i], b1[i], acc00, 0, 0, 0);
                    acc10 = __builtin_amdgcn_mfma_f32_16x16x4f32(b2[i], b2[i], acc10, 0, 0, 0);
                } else {
                    acc00 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[i], b0[i], acc00, 0, 0, 0);
                }
            }
        };
        load(p0, p1, p2, 0);
#pragma unroll
        for (int g = 0; g < NQ; g += 2 * GQ) {
            load(q0, q1, q2, g + GQ);
            __builtin_amdgcn_sched_barrier(0);
            gram(p0, p1, p2);
            __builtin_amdgcn_sched_barrier(0);
            if (g + 2 * GQ < NQ) load(p0, p1, p2, g + 2 * GQ);
            __builtin_amdgcn_sched_barrier(0);
            gram(q0, q1, q2);
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();
        /* promote to f64 every 256 rows */
        if (R == 4 || (step & 1)) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                g00[r] += (double)acc00[r];
                acc00[r] = 0.0f;
                if (KC == 32) {
                    g10[r] += (double)acc10[r];
                    acc10[r] = 0.0f;
                }
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        g00[r] += (double)acc00[r];
        if (KC == 32) g10[r] += (double)acc10[r];
    }
    const int K = k + 2;
    double *pg = pgram + (size_t)item * (size_t)tri_stride;
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = 4 * (lane >> 4) + r;
        if (KC == 16 && i >= j && i < K) pg[i * (i + 1) / 2 + j] = g00[r];
        if (KC == 32) ft_pair_store<6>(pg, i, j, K, g00[r], g10[r]);
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

/* Assembly of the plain FP64 arm on the fit VM: the same work items, row
   contract and Gram layout as vm_asm_kernel, with every row quantity in f64.
   Two f64 rows per lane (the vm_d2 pack: its register files have the slot
   counts of the 2-row f32 block, so the streams are compiled for that shape
   with this block's column stride), the value and adjoint sweeps on f64
   rows with f64 constants, the augmented row [s*J | s*r | 0] in an f64 LDS
   block and the Gram on FP64 MFMA (v_mfma_f64_16x16x4f64) accumulated in
   f64 throughout.  The zero r_lo column keeps the control kernel's
   J^T r = J^T r_hi + J^T r_lo exact.  A row is usable when its f64
   prediction and residual are finite (no f32 magnitude rule). */
template <int KC, bool POLISH>
__global__ __launch_bounds__(64) void vm_asm64_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run,
    const double *__restrict__ cp64, double *__restrict__ pgram, int tri_stride,
    double *__restrict__ ploss, int *__restrict__ pbad, int *__restrict__ pskip)
{
    typedef vm_d2 V;
    typedef dv16 FU;
    typedef dv16 FB;
    typedef dv8 FA;
    typedef VmPack<V> P;
    typedef double double4v __attribute__((ext_vector_type(4)));
    constexpr int R = 2;
    constexpr int XS = VM64_XS(R);
    constexpr int KCA = (KC == 32) ? FT_KMAX + 2 : KC;
    static_assert(KC != 32 || KCA == 22, "the KC = 32 column pairing covers 22 columns");
    constexpr int BLOCK = 64 * R;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const FtCand d = cands[cand];
    const FtRows Rw = rowsd[d.rowset];
    const int lane = threadIdx.x;
    const int k = d.k;

    __shared__ double X[KCA * XS];
    for (int i = lane; i < KCA * XS; i += 64) X[i] = 0.0;
    __syncthreads();

    const double *cst = cp64 + (size_t)cand * FT_KMAX;
    const uint2 *fc = (const uint2 *)(code + d.fwd_off);
    const uint2 *rc = (const uint2 *)(code + d.rev_off);
    char *xl = (char *)(X + lane);
    double4v acc00 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0};
    double lsum = 0.0;
    int bad = 0, skip = 0;
    FU U0 = (FU)(0.0), U1 = U0;
    FB B0 = (FB)(0.0), B1 = B0;
    FA A0 = (FA)(0.0), A1 = A0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += BLOCK) {
        int q[R];
        bool in[R];
        double wb[R], yv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + 64 * r + lane;
            in[r] = row < row_end;
            q[r] = in[r] ? row : it.row0;
            wb[r] = Rw.w64 ? Rw.w64[q[r]] : 1.0;
            yv[r] = Rw.y64[q[r]];
        }
#pragma unroll
        for (int v = 0; v < VM_MAX_VARS; ++v)
            if (v < Rw.v) {
                const double *col = Rw.cols64 + (size_t)v * (size_t)Rw.n;
                P::set(U0, U1, U1, U1, v, (V){col[q[0]], col[q[1]]});
            }
        const V pred = vm_forward<V, true>(fc, d.m, cst, U0, U1, U1, U1, U0, U1, U1, U1, B0, B1,
                                           B1, B1);
        bool live[R], ok[R];
        double rs[R], rho[R];
        V seed;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            live[r] = in[r] && wb[r] > 0.0;
            ok[r] = live[r] && isfinite(pred[r]) && isfinite(pred[r] - yv[r]);
            const RcRowFit f = rc_row_fit(wb[r], pred[r], yv[r], d.metric);
            seed[r] = ok[r] ? f.s : 0.0;
            rs[r] = f.s * f.r;
            rho[r] = f.rho;
        }
        const V fin = vm_reverse<V, FU, FB, FA>(rc, d.nrev, seed, cst, U0, U1, B0, B1, A0, A1, xl);
        bool fix = false;
        bool use[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool gf = fin[r] == 0.0;
            use[r] = ok[r] && gf;
            if (POLISH) {
                if (live[r] && !ok[r]) bad = 1;
                skip += (ok[r] && !gf) ? 1 : 0;
                if (ok[r]) lsum += wb[r] * rho[r];
            } else {
                if (live[r] && !use[r]) bad = 1;
                if (use[r]) lsum += wb[r] * rho[r];
            }
            fix = fix || (ok[r] && !gf);
        }
        if (__any(fix)) {
            for (int j = 0; j < k; ++j)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (!use[r]) X[j * XS + 64 * r + lane] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) X[k * XS + 64 * r + lane] = use[r] ? rs[r] : 0.0;
        __syncthreads();
        const int col = lane & 15;
        const int sub = lane >> 4;
        const int colc = col < 8 ? col : (8 + col < KCA ? 8 + col : KCA - 1);
        const double *x0 = X + col * XS + sub;
        const double *x1 = X + (6 + col) * XS + sub;
        const double *x2 = X + colc * XS + sub;
#pragma unroll 8
        for (int g = 0; g < BLOCK / 4; ++g) {
            const double a0 = x0[4 * g];
            if (KC == 32) {
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, x1[4 * g], acc00, 0, 0, 0);
                const double a2 = x2[4 * g];
                acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, a2, acc10, 0, 0, 0);
            } else {
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc00, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    /* FP64 MFMA output: lane l, register r holds D[l/16 + 4 r][l % 16] of
       each 16x16 block (not the FP32 form's 4 (l/16) + r). */
    const int K = k + 2;
    double *pg = pgram + (size_t)item * (size_t)tri_stride;
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = (lane >> 4) + 4 * r;
        if (KC == 16 && i >= j && i < K) pg[i * (i + 1) / 2 + j] = acc00[r];
        if (KC == 32) ft_pair_store<6>(pg, i, j, K, acc00[r], acc10[r]);
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

/* Trial / exact loss on the fit VM: the value-pass stream (inputs in U
   slots 0..v-1, no tape), 8 f32 rows or 4 f64 rows per lane, with an 8-slot
   U file of four 64-bit-element vectors. */
template <typename V>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void vm_trial_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const FtCand *__restrict__ cands, const FtRows *__restrict__ rowsd,
    const uint32_t *__restrict__ code, const int *__restrict__ run,
    const typename VmScalar<V>::T *__restrict__ cp, double *__restrict__ ploss,
    double *__restrict__ p64out)
{
    typedef typename VmScalar<V>::T T;
    typedef VmPack<V> P;
    typedef typename P::F8 F;
    constexpr int RR = VmScalar<V>::R;
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
    const uint2 *fc = (const uint2 *)(code + d.t_off);
    const T *cols;
    if (sizeof(T) == 4) cols = (const T *)(const void *)R_.cols;
    else cols = (const T *)(const void *)R_.cols64;
    F U0 = (F)(0), U1 = U0, U2 = U0, U3 = U0, Bd = U0;
    double lsum = 0.0;
    const bool sq_unit = d.metric == RC_LOSS_SQUARED && R_.w64 == nullptr;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64 * RR) {
        int xrow[RR];
        bool live[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            const int row = base + r * 64 + lane;
            live[r] = row < row_end;
            xrow[r] = live[r] ? row : it.row0;
        }
#pragma unroll
        for (int v = 0; v < VM_MAX_VARS; ++v)
            if (v < R_.v) {
                const T *col = cols + (size_t)v * (size_t)R_.n;
                V xv;
#pragma unroll
                for (int r = 0; r < RR; ++r) xv[r] = col[xrow[r]];
                P::set(U0, U1, U2, U3, v, xv);
            }
        const V pred = vm_forward<V, false>(fc, d.t_m, cst, U0, U1, U2, U3, U0, U1, U2, U3, Bd, Bd,
                                            Bd, Bd);
        if (sizeof(T) == 8 && p64out) {
            double *pv = p64out + (size_t)it.p64_blk * 64 - it.row0;
#pragma unroll
            for (int r = 0; r < RR; ++r)
                if (live[r]) pv[xrow[r]] = (double)pred[r];
        }
        /* All of the step's weights and targets are requested at once (one
           wait), and a row that does not count adds +0.0 (the sum is never
           -0), so the sum is the same as skipping it. */
        double wv[RR], yv[RR];
#pragma unroll
        for (int r = 0; r < RR; ++r) {
            wv[r] = R_.w64 ? R_.w64[xrow[r]] : 1.0;
            yv[r] = R_.y64[xrow[r]];
        }
        if (sq_unit) {
            /* Squared loss, unit weights: the general form below with
               wb = 1 (1 * t = t exactly); the square is rounded on its own
               as there. */
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const double rr = (double)pred[r] - yv[r];
                bool hard = !isfinite(pred[r]) || !isfinite(rr);
                if (sizeof(T) == 4) hard = hard || fabs((double)pred[r]) > 1e30;
                const double term = hard ? 1e30 : __dmul_rn(rr, rr);
                lsum += live[r] ? term : 0.0;
            }
        } else {
#pragma unroll
            for (int r = 0; r < RR; ++r) {
                const double wb = wv[r];
                const double rr = (double)pred[r] - yv[r];
                bool hard = !isfinite(pred[r]) || !isfinite(rr);
                if (sizeof(T) == 4) hard = hard || fabs((double)pred[r]) > 1e30;
                const double term = hard ? 1e30 : wb * rc_row_stat((double)pred[r], yv[r], d.metric);
                lsum += (live[r] && wb > 0.0) ? term : 0.0;
            }
        }
    }
    const double ls = rc_wave_sum_f64(lsum);
    if (lane == 0) ploss[item] = ls;
}

} /* namespace */

void vm_launch_asm(int kc, int r, bool polish, int nl, hipStream_t s, const FtItem *items,
                   const int *list, const FtCand *cands, const FtRows *rows, const uint32_t *code,
                   const int *run, const float *cp32, const double *cp64, double *pgram,
                   int tri_stride, double *ploss, int *pbad, int *pskip, const double *p64)
{
    if (nl <= 0) return;
#define VM_ASM(KC, PM, R)                                                                       \
    vm_asm_kernel<KC, PM, R><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp32, cp64, \
                                              pgram, tri_stride, ploss, pbad, pskip, p64)
    if (polish && r == 4 && p64) {
        if (kc == 16) VM_ASM(16, 2, 4);
        else VM_ASM(32, 2, 4);
    } else if (polish && r == 4) {
        if (kc == 16) VM_ASM(16, 1, 4);
        else VM_ASM(32, 1, 4);
    } else if (polish) {
        if (kc == 16) VM_ASM(16, 1, 2);
        else VM_ASM(32, 1, 2);
    } else if (r == 4) {
        if (kc == 16) VM_ASM(16, 0, 4);
        else VM_ASM(32, 0, 4);
    } else {
        if (kc == 16) VM_ASM(16, 0, 2);
        else VM_ASM(32, 0, 2);
    }
#undef VM_ASM
}

void vm_launch_asm64(int kc, bool polish, int nl, hipStream_t s, const FtItem *items,
                     const int *list, const FtCand *cands, const FtRows *rows,
                     const uint32_t *code, const int *run, const double *cp64, double *pgram,
                     int tri_stride, double *ploss, int *pbad, int *pskip)
{
    if (nl <= 0) return;
#define VM_ASM64(KC, P)                                                                        \
    vm_asm64_kernel<KC, P><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp64, pgram, \
                                             tri_stride, ploss, pbad, pskip)
    if (polish) {
        if (kc == 16) VM_ASM64(16, true);
        else VM_ASM64(32, true);
    } else {
        if (kc == 16) VM_ASM64(16, false);
        else VM_ASM64(32, false);
    }
#undef VM_ASM64
}

void vm_launch_trial(bool exact, int nl, hipStream_t s, const FtItem *items, const int *list,
                     const FtCand *cands, const FtRows *rows, const uint32_t *code, const int *run,
                     const float *cp32, const double *cp64, double *ploss, double *p64)
{
    if (nl <= 0) return;
    if (exact)
        vm_trial_kernel<vm_d4><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp64, ploss,
                                                  p64);
    else
        vm_trial_kernel<vm_f8><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp32, ploss,
                                                  nullptr);
}
