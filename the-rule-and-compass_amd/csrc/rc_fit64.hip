/*
 * The plain FP64 device fit (rc_set_fit_mode(ctx, 1)): the denominator of
 * the mixed-precision speed-up.
 *
 * It is the reference device arm's structure with every row pass in f64:
 * one workgroup per candidate runs the whole LM / IRLS loop (fit_kernel,
 * eval.cu:2265-2606), each thread evaluates rows of the post-order records
 * with a value array and an adjoint array (ft_forward_d / ft_backward_d,
 * 1324-1463), a 128-row chunk of scaled gradients is staged in LDS and every
 * normal-equation statistic is owned by one thread that accumulates it over
 * the rows in ascending order (ft_assemble_d, 1493-1630), the trial loss is a
 * fixed-tree f64 reduction (ft_loss_d, 1635-1682) and thread 0 solves
 * (ft_solve, 1874-1949).  The LM schedule, IRLS rounds, Armijo halvings, f64
 * polish and status labels are the same as the mixed-precision arm, so the
 * two arms differ only in arithmetic precision and in how the work is laid
 * out on the machine.
 */
#include "rc_fit_dev.h"
#include "rc_internal.h"

#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int P64_THREADS = 256;
constexpr int P64_CHUNK = 128;
constexpr int P64_STATS = FT_KMAX * (FT_KMAX + 1) / 2 + FT_KMAX + 1; /* 231 */

struct P64Cand {
    uint32_t off;
    int m;
    int k;
    int rowset;
    int metric;
    int pre;
};

__device__ __forceinline__ void p64_forward(const uint32_t *__restrict__ nd, int m,
                                            const FtRows &R, int row, const double *c, double *v)
{
    ft_forward64(nd, m, R.cols64, R.n, row, c, v);
}

__device__ __forceinline__ void p64_backward(const uint32_t *__restrict__ nd, int m, const double *v,
                                             double *adj, double *g, int p)
{
    ft_backward64(nd, m, v, adj, g, p);
}

__device__ __forceinline__ int p64_stat_kind(int s, int tri, int p, int *ei, int *ej)
{
    if (s < tri) {
        int i = 0;
        while ((i + 1) * (i + 2) / 2 <= s) ++i;
        *ei = i;
        *ej = s - i * (i + 1) / 2;
        return 0;
    }
    if (s < tri + p) return 1;
    return 2;
}

struct P64Smem {
    double sc[FT_KMAX];
    double strial[FT_KMAX];
    double sdelta[FT_KMAX];
    double sjtj[FT_KMAX * (FT_KMAX + 1) / 2];
    double sjtr[FT_KMAX];
    double sL[FT_KMAX * FT_KMAX];
    double sdg[P64_CHUNK * FT_KMAX];
    double sdr[P64_CHUNK];
    double sdl[P64_CHUNK];
    double sred[P64_THREADS];
    double sbest, slambda, sloss;
    int sbad, sskip, ssolved, srank, simproved, sstart;
};

/* Assemble J^T W J, J^T W r and sum w*rho over the rows in f64.  `polish`
   selects the skip-bad-gradient semantics of the exact polish. */
__device__ void p64_assemble(P64Smem &S, const uint32_t *nd, int m, const FtRows &R, int p,
                             int metric, const double *c, int polish)
{
    const int tid = threadIdx.x;
    const int tri = p * (p + 1) / 2;
    const int total = tri + p + 1;
    double acc[2] = {0.0, 0.0};
    int own[2], kind[2], oi[2], oj[2];
    int nown = 0;
    for (int s = tid; s < total && nown < 2; s += P64_THREADS) {
        own[nown] = s;
        kind[nown] = p64_stat_kind(s, tri, p, &oi[nown], &oj[nown]);
        ++nown;
    }
    for (int base = 0; base < R.n; base += P64_CHUNK) {
        const int rows = min(P64_CHUNK, R.n - base);
        if (tid < P64_CHUNK) {
            const int row = base + tid;
            const double wr = (row < R.n && R.w64) ? R.w64[row] : 1.0;
            double dgl[FT_KMAX];
            double dr = 0.0, dl = 0.0;
            for (int j = 0; j < p; ++j) dgl[j] = 0.0;
            if (row < R.n && wr > 0.0) {
                double v[FT_MAX_NODES], adj[FT_MAX_NODES], g[FT_KMAX];
                p64_forward(nd, m, R, row, c, v);
                const double pred = v[m - 1];
                const double yv = R.y64[row];
                const bool ok = isfinite(pred) && isfinite(pred - yv);
                p64_backward(nd, m, v, adj, g, p);
                bool gok = true;
                for (int j = 0; j < p; ++j) if (!isfinite(g[j])) gok = false;
                const RcRowFit rf = rc_row_fit(wr, pred, yv, metric);
                const double r = rf.r, s = rf.s;
                const bool use = ok && gok;
                for (int j = 0; j < p; ++j) dgl[j] = use ? s * g[j] : 0.0;
                dr = use ? s * r : 0.0;
                dl = (polish ? ok : use) ? wr * rf.rho : 0.0;
                if (polish) {
                    if (!ok) S.sbad = 1;
                    if (ok && !gok) atomicAdd(&S.sskip, 1);
                } else if (!use) {
                    S.sbad = 1;
                }
            }
            for (int j = 0; j < p; ++j) S.sdg[tid * FT_KMAX + j] = dgl[j];
            S.sdr[tid] = dr;
            S.sdl[tid] = dl;
        }
        __syncthreads();
        for (int o = 0; o < nown; ++o) {
            if (kind[o] == 0) {
                for (int q = 0; q < rows; ++q) acc[o] += S.sdg[q * FT_KMAX + oi[o]] * S.sdg[q * FT_KMAX + oj[o]];
            } else if (kind[o] == 1) {
                const int wj = own[o] - tri;
                for (int q = 0; q < rows; ++q) acc[o] += S.sdg[q * FT_KMAX + wj] * S.sdr[q];
            } else {
                for (int q = 0; q < rows; ++q) acc[o] += S.sdl[q];
            }
        }
        __syncthreads();
    }
    for (int o = 0; o < nown; ++o) {
        if (kind[o] == 0) S.sjtj[own[o]] = acc[o];
        else if (kind[o] == 1) S.sjtr[own[o] - tri] = acc[o];
        else S.sloss = acc[o];
    }
    __syncthreads();
}

/* Fixed-tree f64 reduction of sum w*rho at constants c. */
__device__ double p64_loss(P64Smem &S, const uint32_t *nd, int m, const FtRows &R, int metric,
                           const double *c)
{
    const int tid = threadIdx.x;
    double local = 0.0;
    for (int row = tid; row < R.n; row += P64_THREADS) {
        const double wr = R.w64 ? R.w64[row] : 1.0;
        if (wr <= 0.0) continue;
        double v[FT_MAX_NODES];
        p64_forward(nd, m, R, row, c, v);
        const double pred = v[m - 1];
        const double yv = R.y64[row];
        local += (!isfinite(pred) || !isfinite(pred - yv)) ? 1e30 : wr * rc_row_stat(pred, yv, metric);
    }
    S.sred[tid] = local;
    __syncthreads();
    for (int st = P64_THREADS / 2; st > 0; st >>= 1) {
        if (tid < st) S.sred[tid] += S.sred[tid + st];
        __syncthreads();
    }
    const double total = S.sred[0];
    __syncthreads();
    return total;
}

__global__ __launch_bounds__(P64_THREADS) void p64_fit_kernel(
    const uint32_t *__restrict__ prog, const P64Cand *__restrict__ cands,
    const FtRows *__restrict__ rowsd, const double *__restrict__ c0, int kmax, int iters,
    double *__restrict__ outc, double *__restrict__ outsse, int *__restrict__ outstatus)
{
    extern __shared__ char p64_raw[];
    P64Smem &S = *reinterpret_cast<P64Smem *>(p64_raw);
    const int cand = blockIdx.x;
    const int tid = threadIdx.x;
    const P64Cand d = cands[cand];
    const int p = d.k;
    if (d.pre >= 0) {
        if (tid == 0) {
            for (int j = 0; j < kmax; ++j) outc[(size_t)cand * kmax + j] = c0[(size_t)cand * kmax + j];
            outsse[cand] = d.pre == 7 ? 0.0 : __builtin_nan("");
            outstatus[cand] = d.pre;
        }
        return;
    }
    const FtRows R = rowsd[d.rowset];
    const uint32_t *nd = prog + d.off;
    const int m = d.m;
    const int metric = d.metric;
    if (tid < p) S.sc[tid] = c0[(size_t)cand * kmax + tid];
    if (tid == 0) {
        S.sbad = 0; S.srank = 0; S.sskip = 0; S.simproved = 0; S.sstart = 0;
        S.slambda = 1e-3; S.sbest = 0.0;
    }
    __syncthreads();
    const int irls = (metric == RC_LOSS_ABSOLUTE);
    const int rounds = irls ? (iters < 1 ? 1 : (iters < 4 ? iters : 4)) : 1;
    const int inner = irls ? (iters / rounds < 1 ? 1 : iters / rounds) : iters;
    int status = 0;
    double best = 0.0;
    if (iters > 0) {
        p64_assemble(S, nd, m, R, p, metric, S.sc, 0);
        int assembled = 1;
        best = S.sloss;
        if (tid == 0) S.sstart = S.sbad ? 0 : 1;
        __syncthreads();
        if (S.sbad) status = 2;
        for (int round = 0; round < rounds && status == 0 && best != 0.0; ++round) {
            if (irls && round > 0) assembled = 0;
            for (int it = 0; it < inner && status == 0; ++it) {
                if (!assembled) {
                    p64_assemble(S, nd, m, R, p, metric, S.sc, 0);
                    assembled = 1;
                    if (S.sbad) { status = 2; break; }
                }
                if (tid == 0) {
                    const int rc = ft_solve(S.sjtj, S.sjtr, S.slambda, p, S.sdelta, S.sL);
                    S.ssolved = (rc != 1);
                    if (rc == 4) S.srank = 1;
                }
                __syncthreads();
                if (!S.ssolved) { status = 1; break; }
                if (tid < p) S.strial[tid] = fmin(1e6, fmax(-1e6, S.sc[tid] + S.sdelta[tid]));
                __syncthreads();
                double tsse = p64_loss(S, nd, m, R, metric, S.strial);
                for (int half = 0; half < (irls ? 2 : 0) && !(tsse < best); ++half) {
                    if (tid < p) S.strial[tid] = S.sc[tid] + 0.5 * (S.strial[tid] - S.sc[tid]);
                    __syncthreads();
                    tsse = p64_loss(S, nd, m, R, metric, S.strial);
                }
                if (tsse < best) {
                    if (tid < p) S.sc[tid] = S.strial[tid];
                    best = tsse;
                    if (tid == 0) { S.slambda = fmax(S.slambda / 8.0, 1e-12); S.simproved = 1; }
                    assembled = 0;
                    __syncthreads();
                    if (best == 0.0) break;
                } else {
                    if (tid == 0) S.slambda *= 6.0;
                    __syncthreads();
                    if (S.slambda > 1e9) break;
                }
                __syncthreads();
            }
        }
        if (status == 0) {
            const double exact = p64_loss(S, nd, m, R, metric, S.sc);
            if (isfinite(exact)) best = exact;
        }
        for (int it = 0; it < 4 && status == 0 && best != 0.0; ++it) {
            p64_assemble(S, nd, m, R, p, metric, S.sc, 1);
            if (S.sbad) { status = 2; break; }
            if (tid == 0) {
                const int rc = ft_solve(S.sjtj, S.sjtr, 0.0, p, S.sdelta, S.sL);
                S.ssolved = (rc != 1);
                if (rc == 4) S.srank = 1;
            }
            __syncthreads();
            if (!S.ssolved) break;
            if (tid < p) S.strial[tid] = fmin(1e6, fmax(-1e6, S.sc[tid] + S.sdelta[tid]));
            __syncthreads();
            double tsse = p64_loss(S, nd, m, R, metric, S.strial);
            if (irls && !(tsse < best)) {
                if (tid < p) S.strial[tid] = S.sc[tid] + 0.5 * (S.strial[tid] - S.sc[tid]);
                __syncthreads();
                tsse = p64_loss(S, nd, m, R, metric, S.strial);
            }
            if (tsse < best) {
                if (tid < p) S.sc[tid] = S.strial[tid];
                best = tsse;
                if (tid == 0) S.simproved = 1;
                __syncthreads();
            } else {
                break;
            }
        }
    } else {
        best = p64_loss(S, nd, m, R, metric, S.sc);
    }
    __syncthreads();
    if (tid == 0) {
        if (status == 1 || status == 2) {
            status = !S.sstart ? 7 : (S.simproved ? 5 : 6);
        } else if (status == 0) {
            if (S.sskip > 0) status = 8;
            else if (S.srank) status = 4;
            else if (!S.simproved && iters > 0) status = 6;
        }
        for (int j = 0; j < kmax; ++j)
            outc[(size_t)cand * kmax + j] = j < p ? S.sc[j] : c0[(size_t)cand * kmax + j];
        outsse[cand] = best;
        outstatus[cand] = status;
    }
}

} /* namespace */

/* Host driver: pre[] holds host-decided statuses (-1 = fit on the device). */
int rc_fit64_run(RcCtx *c, const uint32_t *prog, const uint32_t *off, const int *pcount, int nprog,
                 int kmax, const double *c0, const std::vector<int> &rowset,
                 const std::vector<int> &metric, const std::vector<int> &pre, int iters,
                 double *outc, double *outsse, int *outstatus)
{
    std::vector<P64Cand> cands((size_t)nprog);
    for (int i = 0; i < nprog; ++i) {
        cands[i].off = off[i];
        cands[i].m = (int)(off[i + 1] - off[i]) / 4;
        cands[i].k = pre[i] >= 0 ? 0 : pcount[i];
        /* the logistic arm reads the label signs in the target slot */
        cands[i].rowset = metric[i] == RC_LOSS_LOGISTIC ? c->nsets : rowset[i];
        cands[i].metric = metric[i];
        cands[i].pre = pre[i];
    }
    const std::vector<FtRows> rowsd = rc_fit_rows(c);
    const size_t words = off[nprog];
    uint32_t *dp = nullptr;
    P64Cand *dc = nullptr;
    FtRows *dr = nullptr;
    double *d0 = nullptr, *doc = nullptr, *dsse = nullptr;
    int *dst = nullptr;
    int rc = -1;
    hipStream_t s = c->s_fit;
    const size_t smem = sizeof(P64Smem);
    if (hipMalloc((void **)&dp, std::max<size_t>(words, 1) * 4) == hipSuccess &&
        hipMalloc((void **)&dc, cands.size() * sizeof(P64Cand)) == hipSuccess &&
        hipMalloc((void **)&dr, rowsd.size() * sizeof(FtRows)) == hipSuccess &&
        hipMalloc((void **)&d0, (size_t)nprog * kmax * 8) == hipSuccess &&
        hipMalloc((void **)&doc, (size_t)nprog * kmax * 8) == hipSuccess &&
        hipMalloc((void **)&dsse, (size_t)nprog * 8) == hipSuccess &&
        hipMalloc((void **)&dst, (size_t)nprog * 4) == hipSuccess) {
        (void)hipMemcpy(dp, prog, words * 4, hipMemcpyHostToDevice);
        (void)hipMemcpy(dc, cands.data(), cands.size() * sizeof(P64Cand), hipMemcpyHostToDevice);
        (void)hipMemcpy(dr, rowsd.data(), rowsd.size() * sizeof(FtRows), hipMemcpyHostToDevice);
        (void)hipMemcpy(d0, c0, (size_t)nprog * kmax * 8, hipMemcpyHostToDevice);
        (void)hipFuncSetAttribute((const void *)p64_fit_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)smem);
        (void)hipEventRecord(c->t0, s);
        p64_fit_kernel<<<nprog, P64_THREADS, smem, s>>>(dp, dc, dr, d0, kmax, iters, doc, dsse, dst);
        (void)hipEventRecord(c->t1, s);
        if (hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess) {
            (void)hipMemcpy(outc, doc, (size_t)nprog * kmax * 8, hipMemcpyDeviceToHost);
            (void)hipMemcpy(outsse, dsse, (size_t)nprog * 8, hipMemcpyDeviceToHost);
            (void)hipMemcpy(outstatus, dst, (size_t)nprog * 4, hipMemcpyDeviceToHost);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, c->t0, c->t1);
            c->stats.kernel_ms = ms;
            rc = 0;
        }
    }
    (void)hipGetLastError();
    if (dp) (void)hipFree(dp);
    if (dc) (void)hipFree(dc);
    if (dr) (void)hipFree(dr);
    if (d0) (void)hipFree(d0);
    if (doc) (void)hipFree(doc);
    if (dsse) (void)hipFree(dsse);
    if (dst) (void)hipFree(dst);
    return rc;
}
