/*
 * Row kernels of the correlation fit arm (rc_fit_batch metrics 2 and 3),
 * compiled with the fit VM kernels' unstructured uniform regions.
 *
 * Reference behaviour: ft_corr_assemble and ft_corr_score
 * (eval.cu:1970-2239); the CPU arm is fit/correlation.rs.  The objective is
 * 1 - |rho| (or 1 - rho) of the weighted Pearson correlation of the
 * prediction f and the target y, and every moment its gradient and
 * Gauss-Newton curvature use is an entry of the weighted Gram matrix of
 * z = [g_1 .. g_p, 1, f, y] with g = df/dc.
 *
 * MI355X layout: one wave per work item (candidate, row tile), like the
 * other fit arms.  The assembly writes sqrt(w) z for 128 f64 rows (two per
 * lane) into an LDS block and reduces it on FP64 MFMA
 * (v_mfma_f64_16x16x4f64) into the item's Gram triangle; the score pass
 * sums the six moments of one constant vector per item.  Programs that fit
 * the fit VM run its f64 streams (vm_d2 value and adjoint sweeps, vm_d4
 * value pass); the others run the post-order record sweeps, one row per
 * lane.  The control kernel (rc_corr.hip) reduces the item partials in tile
 * order, so a candidate's result does not depend on its batch.
 */
#include "rc_corr.h"
#include "rc_fit_dev.h"
#include "rc_fitvm_dev.h"
#include "rc_fitvm_kern.h"

#include <math.h>

namespace {

typedef double cr_d4 __attribute__((ext_vector_type(4)));

/* A row's contribution is usable: finite prediction within the f32 hard-fail
   magnitude (the reference arm's rows are f32) and a finite target. */
__device__ __forceinline__ bool cr_row_ok(double f, double y)
{
    return isfinite(f) && fabs(f) <= 1e30 && isfinite(y);
}

template <int KC, bool VM>
__global__ __launch_bounds__(64) void cr_asm_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const CrCand *__restrict__ cands, FtRows Rw, const uint32_t *__restrict__ code,
    const int *__restrict__ run, const double *__restrict__ cp, double *__restrict__ pgram,
    int *__restrict__ pbad)
{
    typedef vm_d2 V;
    typedef VmPack<V> P;
    constexpr int R = VM ? 2 : 1;
    constexpr int XS = VM64_XS(R);
    constexpr int KCA = (KC == 32) ? CR_ZMAX : KC;
    constexpr int PS = KCA - 16; /* the column pairing's shift (ft_pair_store) */
    static_assert(KC != 32 || (PS >= 0 && PS <= 8), "the column pairing covers KCA columns");
    constexpr int BLOCK = 64 * R;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const CrCand d = cands[cand];
    const int lane = threadIdx.x;
    const int k = d.k;
    const int q = k + 3;

    __shared__ double X[KCA * XS];
    for (int i = lane; i < KCA * XS; i += 64) X[i] = 0.0;
    __syncthreads();

    const double *cst = cp + (size_t)cand * FT_KMAX;
    /* KC = 16: acc00 is the Gram block.  KC = 32: acc00 and acc10 are the
       products P and Q of ft_pair_store<PS>. */
    cr_d4 acc00 = {0, 0, 0, 0}, acc10 = {0, 0, 0, 0};
    int bad = 0;
    dv16 U0 = (dv16)(0.0), U1 = U0, B0 = U0, B1 = U0;
    dv8 A0 = (dv8)(0.0), A1 = A0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += BLOCK) {
        int rq[R];
        bool in[R];
        double wb[R], yv[R], f[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + 64 * r + lane;
            in[r] = row < row_end;
            rq[r] = in[r] ? row : it.row0;
            wb[r] = Rw.w64 ? Rw.w64[rq[r]] : 1.0;
            yv[r] = Rw.y64[rq[r]];
        }
        if constexpr (VM) {
            for (int v = 0; v < VM_MAX_VARS; ++v)
                if (v < Rw.v) {
                    const double *col = Rw.cols64 + (size_t)v * (size_t)Rw.n;
                    P::set(U0, U1, U1, U1, v, (V){col[rq[0]], col[rq[R - 1]]});
                }
            const V pred = vm_forward<V, true>((const uint2 *)(code + d.fwd_off), d.m, cst, U0, U1,
                                               U1, U1, U0, U1, U1, U1, B0, B1, B1, B1);
            V seed;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                f[r] = pred[r];
                ok[r] = in[r] && wb[r] > 0.0 && cr_row_ok(f[r], yv[r]);
                seed[r] = ok[r] ? sqrt(wb[r]) : 0.0;
            }
            const V fin = vm_reverse<V, dv16, dv16, dv8>((const uint2 *)(code + d.rev_off), d.nrev,
                                                         seed, cst, U0, U1, B0, B1, A0, A1,
                                                         (char *)(X + lane));
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (in[r] && wb[r] > 0.0 && !(ok[r] && fin[r] == 0.0)) bad = 1;
                ok[r] = ok[r] && fin[r] == 0.0;
            }
        } else {
            /* Record sweeps: one row per lane, gradient straight into the
               block (the weight applied as sqrt(w) like the VM's seed). */
            const uint32_t *recs = code + d.fwd_off;
            double val[FT_MAX_NODES], adj[FT_MAX_NODES], g[FT_KMAX];
            ft_rec_values(recs, d.m, Rw.cols64, Rw.n, rq[0], cst, val);
            f[0] = val[d.m - 1];
            ft_rec_gradient(recs, d.m, val, adj, g, k);
            bool gfin = true;
            for (int j = 0; j < k; ++j) gfin = gfin && isfinite(g[j]);
            const bool live = in[0] && wb[0] > 0.0;
            ok[0] = live && cr_row_ok(f[0], yv[0]) && gfin;
            if (live && !ok[0]) bad = 1;
            const double sw = ok[0] ? sqrt(wb[0]) : 0.0;
            for (int j = 0; j < k; ++j) X[j * XS + lane] = ok[0] ? sw * g[j] : 0.0;
        }
        bool fix = false;
#pragma unroll
        for (int r = 0; r < R; ++r) fix = fix || !ok[r];
        if (VM && __any(fix)) {
            for (int j = 0; j < k; ++j)
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (!ok[r]) X[j * XS + 64 * r + lane] = 0.0;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double sw = ok[r] ? sqrt(wb[r]) : 0.0;
            X[k * XS + 64 * r + lane] = sw;
            X[(k + 1) * XS + 64 * r + lane] = ok[r] ? sw * f[r] : 0.0;
            X[(k + 2) * XS + 64 * r + lane] = ok[r] ? sw * yv[r] : 0.0;
        }
        __syncthreads();
        const int col = lane & 15;
        const int sub = lane >> 4;
        const int colc = col < 8 ? col : (8 + col < KCA ? 8 + col : KCA - 1);
        const double *x0 = X + col * XS + sub;
        const double *x1 = X + (KC == 32 ? PS + col : col) * XS + sub;
        const double *x2 = X + colc * XS + sub;
#pragma unroll 8
        for (int gq = 0; gq < BLOCK / 4; ++gq) {
            const double a0 = x0[4 * gq];
            if (KC == 32) {
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, x1[4 * gq], acc00, 0, 0, 0);
                const double a2 = x2[4 * gq];
                acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, a2, acc10, 0, 0, 0);
            } else {
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, a0, acc00, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    /* FP64 MFMA output: lane l, register r holds D[l/16 + 4 r][l % 16] of
       each 16x16 block. */
    double *pg = pgram + (size_t)item * CR_TRI;
    const int j = lane & 15;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int i = (lane >> 4) + 4 * r;
        if (KC == 16 && i >= j && i < q) pg[i * (i + 1) / 2 + j] = acc00[r];
        if (KC == 32) ft_pair_store<PS>(pg, i, j, q, acc00[r], acc10[r]);
    }
    const int anybad = __any(bad) ? 1 : 0;
    if (lane == 0) pbad[item] = anybad;
}

template <bool VM>
__global__ __launch_bounds__(64) void cr_score_kernel(
    const FtItem *__restrict__ items, const int *__restrict__ list, int nlist,
    const CrCand *__restrict__ cands, FtRows Rw, const uint32_t *__restrict__ code,
    const int *__restrict__ run, const double *__restrict__ cp, double *__restrict__ pmom,
    int *__restrict__ pbad)
{
    typedef vm_d4 V;
    typedef VmPack<V> P;
    constexpr int R = VM ? 4 : 1;
    const int li = blockIdx.x;
    if (li >= nlist) return;
    const int item = list[li];
    const FtItem it = items[item];
    const int cand = it.cand;
    if (!run[cand]) return;
    const CrCand d = cands[cand];
    const int lane = threadIdx.x;
    const double *cst = cp + (size_t)cand * FT_KMAX;
    double mo[CR_NMOM] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int bad = 0;
    dv8 U0 = (dv8)(0.0), U1 = U0, U2 = U0, U3 = U0, Bd = U0;
    const int row_end = it.row0 + it.rows;
    for (int base = it.row0; base < row_end; base += 64 * R) {
        int rq[R];
        bool in[R];
        double f[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int row = base + 64 * r + lane;
            in[r] = row < row_end;
            rq[r] = in[r] ? row : it.row0;
        }
        if constexpr (VM) {
            for (int v = 0; v < VM_MAX_VARS; ++v)
                if (v < Rw.v) {
                    const double *col = Rw.cols64 + (size_t)v * (size_t)Rw.n;
                    V xv;
#pragma unroll
                    for (int r = 0; r < R; ++r) xv[r] = col[rq[r]];
                    P::set(U0, U1, U2, U3, v, xv);
                }
            const V pred = vm_forward<V, false>((const uint2 *)(code + d.t_off), d.t_m, cst, U0, U1,
                                                U2, U3, U0, U1, U2, U3, Bd, Bd, Bd, Bd);
#pragma unroll
            for (int r = 0; r < R; ++r) f[r] = pred[r];
        } else {
            double val[FT_MAX_NODES];
            ft_rec_values(code + d.t_off, d.t_m, Rw.cols64, Rw.n, rq[0], cst, val);
            f[0] = val[d.t_m - 1];
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double wb = Rw.w64 ? Rw.w64[rq[r]] : 1.0;
            const double y = Rw.y64[rq[r]];
            if (!(in[r] && wb > 0.0)) continue;
            if (!cr_row_ok(f[r], y)) {
                bad = 1;
                continue;
            }
            mo[0] += wb;
            mo[1] += wb * f[r];
            mo[2] += wb * f[r] * f[r];
            mo[3] += wb * f[r] * y;
            mo[4] += wb * y;
            mo[5] += wb * y * y;
        }
    }
    double out[CR_NMOM];
#pragma unroll
    for (int qd = 0; qd < CR_NMOM; ++qd) out[qd] = rc_wave_sum_f64(mo[qd]);
    const int anybad = __any(bad) ? 1 : 0;
    if (lane == 0) {
#pragma unroll
        for (int qd = 0; qd < CR_NMOM; ++qd) pmom[(size_t)item * CR_NMOM + qd] = out[qd];
        pbad[item] = anybad;
    }
}

} /* namespace */

void cr_launch_asm(int kc, bool vm, int nl, hipStream_t s, const FtItem *items, const int *list,
                   const CrCand *cands, FtRows rows, const uint32_t *code, const int *run,
                   const double *cp, double *pgram, int *pbad)
{
    if (nl <= 0) return;
#define CR_ASM(KC, VMV)                                                                         \
    cr_asm_kernel<KC, VMV><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp, pgram, pbad)
    if (vm) {
        if (kc == 16) CR_ASM(16, true);
        else CR_ASM(32, true);
    } else {
        if (kc == 16) CR_ASM(16, false);
        else CR_ASM(32, false);
    }
#undef CR_ASM
}

void cr_launch_score(bool vm, int nl, hipStream_t s, const FtItem *items, const int *list,
                     const CrCand *cands, FtRows rows, const uint32_t *code, const int *run,
                     const double *cp, double *pmom, int *pbad)
{
    if (nl <= 0) return;
    if (vm)
        cr_score_kernel<true><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp, pmom, pbad);
    else
        cr_score_kernel<false><<<nl, 64, 0, s>>>(items, list, nl, cands, rows, code, run, cp, pmom, pbad);
}
