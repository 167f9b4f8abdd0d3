/*
 * The correlation fit arm of rc_fit_batch (metric 2: minimise 1 - |rho|,
 * metric 3: 1 - rho) with the affine gauge freeze, on gfx950.
 *
 * Reference behaviour: the correlation branch of fit_kernel
 * (eval.cu:2320-2448) with ft_corr_assemble, ft_corr_score and
 * ft_corr_mask_gauge (1970-2262); the CPU arm it mirrors is
 * fit/correlation.rs (moments, the Gauss-Newton curvature of rho, the gauge
 * renormalisation a -> sign(a), b -> 0).
 *
 * rho is the weighted Pearson correlation of the prediction f and the target
 * y.  With g = d f / d c, the assembly needs the weighted Gram matrix of the
 * row vector z = [g_1 .. g_p, 1, f, y]: its entries are every moment the
 * gradient and the curvature use (C_jk, sum w g_j, sum w g_j f, sum w g_j y,
 * W, sum w f, sum w f^2, sum w f y, sum w y, sum w y^2).  So one owner-thread
 * Gram of (p + 3) columns per 128-row chunk replaces the reference's four
 * statistic kinds, and the centring and the rho formulas run once on one
 * lane.
 *
 * MI355X layout: one workgroup (256 threads) per candidate runs the whole LM
 * loop in f64 (rows, moments and solve), like the plain FP64 arm; a chunk's
 * z rows sit in LDS and every Gram entry is owned by one thread that sums it
 * over the rows in ascending order, so the result is deterministic.  The
 * reference accumulates f32 row values in f64; here the rows are f64 too.
 * A trial whose rows are not all finite scores 1e30, so the LM rejects it.
 */
#include "rc_fit_dev.h"
#include "rc_internal.h"

#include <algorithm>
#include <vector>

namespace {

constexpr int CR_THREADS = 256;
constexpr int CR_CHUNK = 128;
constexpr int CR_Z = FT_KMAX + 3;
constexpr int CR_STATS = CR_Z * (CR_Z + 1) / 2;
static_assert(CR_STATS <= 2 * CR_THREADS, "two Gram entries per thread at most");

struct CrCand {
    uint32_t off;
    int m;
    int k;
    int pre; /* >= 0: a status decided on the host, the kernel copies c0 */
};

struct CrSmem {
    double sc[FT_KMAX];
    double strial[FT_KMAX];
    double sdelta[FT_KMAX];
    double ajtj[FT_KMAX * (FT_KMAX + 1) / 2]; /* assembled curvature and rho gradient */
    double ajtr[FT_KMAX];
    double sjtj[FT_KMAX * (FT_KMAX + 1) / 2]; /* the masked system of one step */
    double sjtr[FT_KMAX];
    double sL[FT_KMAX * FT_KMAX];
    double gram[CR_STATS];
    double sz[CR_CHUNK * CR_Z];
    double sw[CR_CHUNK];
    double sred[CR_THREADS];
    double sscore, svar, syy, srho;
    int sbad, ssolved, srank, simproved, sstart, sflag;
};

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

/* Weighted Gram of z = [g | 1 | f | y] over the rows, then the moments, the
   rho gradient (in ajtr) and its Gauss-Newton curvature (in ajtj), the
   objective and the variances (ft_corr_assemble, eval.cu:2009-2177). */
__device__ void cr_assemble(CrSmem &S, const uint32_t *nd, int m, const FtRows &R, int p,
                            int metric, const double *c)
{
    const int tid = threadIdx.x;
    const int q = p + 3;
    const int total = q * (q + 1) / 2;
    int own[2], oa[2], ob[2];
    double acc[2] = {0.0, 0.0};
    int nown = 0;
    for (int s = tid; s < total && nown < 2; s += CR_THREADS) {
        own[nown] = s;
        ft_tri_index(s, oa[nown], ob[nown]);
        ++nown;
    }
    for (int base = 0; base < R.n; base += CR_CHUNK) {
        const int rows = min(CR_CHUNK, R.n - base);
        if (tid < CR_CHUNK) {
            const int row = base + tid;
            const double wr = (row < R.n && R.w64) ? R.w64[row] : 1.0;
            double *z = S.sz + tid * CR_Z;
            for (int a = 0; a < q; ++a) z[a] = 0.0;
            double wz = 0.0;
            if (row < R.n && wr > 0.0) {
                double v[FT_MAX_NODES], adj[FT_MAX_NODES], g[FT_KMAX];
                ft_forward64(nd, m, R.cols64, R.n, row, c, v);
                const double f = v[m - 1];
                const double t = R.y64[row];
                bool ok = isfinite(f) && fabs(f) <= 1e30 && isfinite(t);
                ft_backward64(nd, m, v, adj, g, p);
                for (int j = 0; j < p; ++j)
                    if (!isfinite(g[j])) ok = false;
                if (ok) {
                    for (int j = 0; j < p; ++j) z[j] = g[j];
                    z[p] = 1.0;
                    z[p + 1] = f;
                    z[p + 2] = t;
                    wz = wr;
                } else {
                    S.sbad = 1;
                }
            }
            S.sw[tid] = wz;
        }
        __syncthreads();
        for (int o = 0; o < nown; ++o) {
            const int a = oa[o], b = ob[o];
            for (int k = 0; k < rows; ++k)
                acc[o] += (S.sw[k] * S.sz[k * CR_Z + a]) * S.sz[k * CR_Z + b];
        }
        __syncthreads();
    }
    for (int o = 0; o < nown; ++o) S.gram[own[o]] = acc[o];
    __syncthreads();
    if (tid == 0) {
        const double W = S.gram[tri(p, p)];
        const double sum_f = S.gram[tri(p + 1, p)];
        const double sum_f2 = S.gram[tri(p + 1, p + 1)];
        const double sum_t = S.gram[tri(p + 2, p)];
        const double sum_ft = S.gram[tri(p + 2, p + 1)];
        const double sum_t2 = S.gram[tri(p + 2, p + 2)];
        double sff = 0.0, syy = 0.0, rho = 0.0;
        for (int i = 0; i < p; ++i)
            for (int j = 0; j <= i; ++j) S.ajtj[tri(i, j)] = S.gram[tri(i, j)];
        for (int j = 0; j < p; ++j) S.ajtr[j] = 0.0;
        if (W > 0.0) {
            const double mf = sum_f / W;
            const double mt = sum_t / W;
            sff = sum_f2 - W * mf * mf;
            syy = sum_t2 - W * mt * mt;
            const double sft = sum_ft - W * mf * mt;
            if (sff > 0.0 && syy > 0.0) rho = sft / sqrt(sff * syy);
            double A[FT_KMAX], B[FT_KMAX], gbar[FT_KMAX];
            for (int j = 0; j < p; ++j) {
                gbar[j] = S.gram[tri(p, j)] / W;
                A[j] = S.gram[tri(p + 1, j)] - W * gbar[j] * mf;
                B[j] = S.gram[tri(p + 2, j)] - W * gbar[j] * mt;
            }
            if (sff > 0.0 && syy > 0.0) {
                const double root = sqrt(sff * syy);
                for (int j = 0; j < p; ++j) S.ajtr[j] = B[j] / root - rho * A[j] / sff;
                for (int i = 0; i < p; ++i)
                    for (int j = 0; j <= i; ++j) {
                        const double centred = S.gram[tri(i, j)] - W * gbar[i] * gbar[j];
                        S.ajtj[tri(i, j)] = (centred - A[i] * A[j] / sff) / sff;
                    }
            } else {
                for (int i = 0; i < p; ++i)
                    for (int j = 0; j <= i; ++j) S.ajtj[tri(i, j)] = (i == j) ? 1.0 : 0.0;
            }
        }
        S.svar = sff;
        S.syy = syy;
        S.srho = rho;
        S.sscore = metric == RC_LOSS_CORRELATION_SIGNED ? 1.0 - rho : 1.0 - fabs(rho);
        if (!isfinite(S.sscore) || !isfinite(sff) || !isfinite(syy)) S.sbad = 1;
    }
    __syncthreads();
}

/* The objective at constants c: six weighted moments in a fixed reduction
   tree (ft_corr_score, eval.cu:2183-2239); a non-finite row scores 1e30. */
__device__ double cr_score(CrSmem &S, const uint32_t *nd, int m, const FtRows &R, int metric,
                           const double *c)
{
    const int tid = threadIdx.x;
    double mo[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int bad = 0;
    for (int row = tid; row < R.n; row += CR_THREADS) {
        const double wr = R.w64 ? R.w64[row] : 1.0;
        if (!(wr > 0.0)) continue;
        double v[FT_MAX_NODES];
        ft_forward64(nd, m, R.cols64, R.n, row, c, v);
        const double f = v[m - 1];
        const double t = R.y64[row];
        if (!isfinite(f) || fabs(f) > 1e30 || !isfinite(t)) {
            bad = 1;
            break;
        }
        mo[0] += wr;
        mo[1] += wr * f;
        mo[2] += wr * f * f;
        mo[3] += wr * f * t;
        mo[4] += wr * t;
        mo[5] += wr * t * t;
    }
    if (tid == 0) S.sflag = 0;
    __syncthreads();
    if (bad) S.sflag = 1;
    for (int qd = 0; qd < 6; ++qd) {
        S.sred[tid] = mo[qd];
        __syncthreads();
        for (int st = CR_THREADS / 2; st > 0; st >>= 1) {
            if (tid < st) S.sred[tid] += S.sred[tid + st];
            __syncthreads();
        }
        mo[qd] = S.sred[0];
        __syncthreads();
    }
    if (S.sflag) return 1e30;
    const double W = mo[0];
    if (!(W > 0.0)) return 1e30;
    const double mf = mo[1] / W;
    const double mt = mo[4] / W;
    const double sff = mo[2] - W * mf * mf;
    const double syy = mo[5] - W * mt * mt;
    const double sft = mo[3] - W * mf * mt;
    if (!(sff > 0.0) || !(syy > 0.0)) return 1.0;
    const double rho = sft / sqrt(sff * syy);
    return metric == RC_LOSS_CORRELATION_SIGNED ? 1.0 - rho : 1.0 - fabs(rho);
}

/* Objective gradient from the rho gradient (1 - sigma rho), and the gauge
   slots pinned: identity row / column and a zero right-hand side, so their
   step is exactly 0 (ft_corr_mask_gauge, eval.cu:2244-2262). */
__device__ void cr_mask_gauge(double *jtj, double *jtr, int p, double rho, int metric, int gs,
                              int go)
{
    const double sigma =
        metric == RC_LOSS_CORRELATION_SIGNED ? 1.0 : (rho >= 0.0 ? 1.0 : -1.0);
    for (int j = 0; j < p; ++j) jtr[j] = -sigma * jtr[j];
    for (int i = 0; i < p; ++i) {
        const bool gi = gs == i || go == i;
        for (int j = 0; j <= i; ++j) {
            const bool gj = gs == j || go == j;
            if (gi || gj) jtj[tri(i, j)] = (i == j) ? 1.0 : 0.0;
        }
        if (gi) jtr[i] = 0.0;
    }
}

/* One damped (or, lambda = 0, undamped) step from sc into strial.  Each
   step masks a fresh copy of the assembled system: the right-hand side is
   sigma times the rho gradient at every step, as the CPU arm builds it
   (correlation.rs fit_lm).  The reference device arm masks its assembled
   arrays in place at every step, so after a rejected step it reuses a
   gradient already negated once. */
__device__ bool cr_step(CrSmem &S, int p, double lambda, int metric, int gs, int go)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        for (int t = 0; t < p * (p + 1) / 2; ++t) S.sjtj[t] = S.ajtj[t];
        for (int j = 0; j < p; ++j) S.sjtr[j] = S.ajtr[j];
        cr_mask_gauge(S.sjtj, S.sjtr, p, S.srho, metric, gs, go);
        const int solve = ft_solve(S.sjtj, S.sjtr, lambda, p, S.sdelta, S.sL);
        S.ssolved = solve != 1;
        if (solve == 4) S.srank = 1;
    }
    __syncthreads();
    if (!S.ssolved) return false;
    if (tid < p) S.strial[tid] = fmin(1e6, fmax(-1e6, S.sc[tid] + S.sdelta[tid]));
    __syncthreads();
    return true;
}

__global__ __launch_bounds__(CR_THREADS) void cr_fit_kernel(
    const uint32_t *__restrict__ prog, const CrCand *__restrict__ cands, FtRows R,
    const double *__restrict__ c0, const int *__restrict__ gauge, int kmax, int iters,
    int metric, double *__restrict__ outc, double *__restrict__ outsse, int *__restrict__ outstatus)
{
    __shared__ CrSmem S;
    const int cand = blockIdx.x;
    const int tid = threadIdx.x;
    const CrCand d = cands[cand];
    if (d.pre >= 0) {
        if (tid == 0) {
            for (int j = 0; j < kmax; ++j) outc[(size_t)cand * kmax + j] = c0[(size_t)cand * kmax + j];
            outsse[cand] = __builtin_nan("");
            outstatus[cand] = d.pre;
        }
        return;
    }
    const int p = d.k;
    const uint32_t *nd = prog + d.off;
    const int m = d.m;
    const int gs = gauge ? gauge[2 * cand] : -1;
    const int go = gauge ? gauge[2 * cand + 1] : -1;
    if (tid < p) S.sc[tid] = c0[(size_t)cand * kmax + tid];
    if (tid == 0) {
        S.sbad = 0;
        S.srank = 0;
        S.simproved = 0;
        S.sstart = 0;
        /* Gauge renormalisation: scale -> sign(scale), offset -> 0. */
        if (gs >= 0 && gs < p && isfinite(S.sc[gs]) && S.sc[gs] != 0.0)
            S.sc[gs] = S.sc[gs] > 0.0 ? 1.0 : -1.0;
        if (go >= 0 && go < p) S.sc[go] = 0.0;
    }
    __syncthreads();
    int status = 0;
    double best = 1.0;
    if (iters > 0) {
        cr_assemble(S, nd, m, R, p, metric, S.sc);
        if (tid == 0) S.sstart = S.sbad ? 0 : 1;
        __syncthreads();
        if (S.sbad) {
            status = 2;
        } else if (!(S.svar > 0.0) || !(S.syy > 0.0)) {
            best = 1.0;
        } else if (S.svar < 1e-24 * S.syy) {
            status = 1;
        } else {
            best = S.sscore;
            double lambda = 1e-3;
            bool assembled = true;
            for (int it = 0; it < iters && status == 0; ++it) {
                if (!assembled) {
                    cr_assemble(S, nd, m, R, p, metric, S.sc);
                    assembled = true;
                    if (S.sbad) {
                        status = 2;
                        break;
                    }
                    if (S.svar < 1e-24 * S.syy) {
                        status = 1;
                        break;
                    }
                }
                if (!cr_step(S, p, lambda, metric, gs, go)) {
                    status = 1;
                    break;
                }
                const double trial = cr_score(S, nd, m, R, metric, S.strial);
                if (trial < best) {
                    if (tid < p) S.sc[tid] = S.strial[tid];
                    best = trial;
                    lambda = fmax(lambda / 8.0, 1e-12);
                    assembled = false;
                    if (tid == 0) S.simproved = 1;
                    __syncthreads();
                    if (best <= 0.0) break;
                } else {
                    lambda *= 6.0;
                    if (lambda > 1e9) break;
                }
            }
            for (int it = 0; it < 4 && status == 0 && best > 0.0; ++it) {
                cr_assemble(S, nd, m, R, p, metric, S.sc);
                if (S.sbad) {
                    status = 2;
                    break;
                }
                if (!cr_step(S, p, 0.0, metric, gs, go)) break;
                const double trial = cr_score(S, nd, m, R, metric, S.strial);
                if (trial < best) {
                    if (tid < p) S.sc[tid] = S.strial[tid];
                    best = trial;
                    if (tid == 0) S.simproved = 1;
                    __syncthreads();
                } else {
                    break;
                }
            }
        }
    } else {
        best = cr_score(S, nd, m, R, metric, S.sc);
    }
    __syncthreads();
    if (tid == 0) {
        /* Labels (eval.cu:2429-2445): a failure after an improvement is 5,
           before one 6, without a usable start 7; a rank-deficient solve on
           the way is 4. */
        if (status == 1 || status == 2) status = !S.sstart ? 7 : (S.simproved ? 5 : 6);
        else if (status == 0 && S.srank) status = 4;
        for (int i = 0; i < kmax; ++i)
            outc[(size_t)cand * kmax + i] = i < p ? S.sc[i] : c0[(size_t)cand * kmax + i];
        outsse[cand] = best;
        outstatus[cand] = status;
    }
}

} /* namespace */

/* The correlation arm of rc_fit_batch: one workgroup per candidate. */
int rc_corr_run(RcCtx *c, const uint32_t *prog, const uint32_t *off, const int *pcount, int nprog,
                int kmax, const double *c0, const int *gauge, int iters, int metric, double *outc,
                double *outsse, int *outstatus)
{
    RcRows *r = &c->sets[0];
    if (r->n <= 0 || rc_rows_ensure_f64(c, r) != 0) return -1;
    std::vector<CrCand> cands((size_t)nprog);
    for (int i = 0; i < nprog; ++i) {
        if (off[i + 1] < off[i]) return -1;
        const uint32_t words = off[i + 1] - off[i];
        if (words == 0 || words % 4 != 0 || words / 4 > FT_MAX_NODES) return -1;
        cands[i].off = off[i];
        cands[i].m = (int)(words / 4);
        cands[i].k = pcount[i];
        cands[i].pre = (pcount[i] < 1 || pcount[i] > kmax) ? 1 : -1;
        if (cands[i].pre >= 0) continue;
        /* Records the kernel can run without a stray read. */
        const uint32_t *nd = prog + off[i];
        for (int q = 0; q < cands[i].m; ++q) {
            const uint32_t op = nd[4 * q], k0 = nd[4 * q + 1], k1 = nd[4 * q + 2], slot = nd[4 * q + 3];
            if (op == OP_CONST) {
                if (slot >= (uint32_t)pcount[i]) return -1;
            } else if (op == OP_VAR) {
                if ((int)slot < 0 || (int)slot >= r->v) return -1;
            } else if (k0 >= (uint32_t)q ||
                       ((op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                         op == OP_POW) && k1 >= (uint32_t)q)) {
                return -1;
            }
        }
    }
    FtRows R;
    R.cols = r->cols;
    R.cols64 = r->cols64;
    R.y64 = r->y64;
    R.w64 = r->w64;
    R.n = r->n;
    R.v = r->v;
    const size_t words = off[nprog];
    uint32_t *dp = nullptr;
    CrCand *dc = nullptr;
    double *d0 = nullptr, *doc = nullptr, *dsse = nullptr;
    int *dg = nullptr, *dst = nullptr;
    int rc = -1;
    hipStream_t s = c->s_fit;
    if (hipMalloc((void **)&dp, std::max<size_t>(words, 1) * 4) == hipSuccess &&
        hipMalloc((void **)&dc, cands.size() * sizeof(CrCand)) == hipSuccess &&
        hipMalloc((void **)&d0, (size_t)nprog * kmax * 8) == hipSuccess &&
        hipMalloc((void **)&doc, (size_t)nprog * kmax * 8) == hipSuccess &&
        hipMalloc((void **)&dsse, (size_t)nprog * 8) == hipSuccess &&
        hipMalloc((void **)&dst, (size_t)nprog * 4) == hipSuccess &&
        (!gauge || hipMalloc((void **)&dg, (size_t)nprog * 2 * 4) == hipSuccess)) {
        bool ok = hipMemcpy(dp, prog, words * 4, hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(dc, cands.data(), cands.size() * sizeof(CrCand),
                            hipMemcpyHostToDevice) == hipSuccess &&
                  hipMemcpy(d0, c0, (size_t)nprog * kmax * 8, hipMemcpyHostToDevice) == hipSuccess &&
                  (!gauge || hipMemcpy(dg, gauge, (size_t)nprog * 2 * 4, hipMemcpyHostToDevice) ==
                                 hipSuccess);
        if (ok) {
            (void)hipEventRecord(c->t0, s);
            cr_fit_kernel<<<nprog, CR_THREADS, 0, s>>>(dp, dc, R, d0, dg, kmax, iters, metric, doc,
                                                      dsse, dst);
            (void)hipEventRecord(c->t1, s);
            if (hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess &&
                hipMemcpy(outc, doc, (size_t)nprog * kmax * 8, hipMemcpyDeviceToHost) == hipSuccess &&
                hipMemcpy(outsse, dsse, (size_t)nprog * 8, hipMemcpyDeviceToHost) == hipSuccess &&
                hipMemcpy(outstatus, dst, (size_t)nprog * 4, hipMemcpyDeviceToHost) == hipSuccess) {
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, c->t0, c->t1);
                c->stats.kernel_ms = ms;
                rc = 0;
            }
        }
    }
    (void)hipGetLastError();
    if (dp) (void)hipFree(dp);
    if (dc) (void)hipFree(dc);
    if (d0) (void)hipFree(d0);
    if (doc) (void)hipFree(doc);
    if (dsse) (void)hipFree(dsse);
    if (dst) (void)hipFree(dst);
    if (dg) (void)hipFree(dg);
    return rc;
}
