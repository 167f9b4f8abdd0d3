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
 * MI355X layout: the fit arms' decomposition.  Each candidate's rows are
 * split into work items (8192-row tiles); every LM slot is a batch-wide
 * moment-block assembly over the items of the candidates that need one, a
 * control launch (one wave per candidate: the item Gram partials reduced in
 * tile order, the rho gradient and curvature, the gauge mask and the wave
 * solve), a batch-wide score pass and a second control launch (acceptance
 * and the lambda schedule).  Rows, moments and solves are f64.  The row
 * kernels are in rc_corr_kern.hip.
 */
#include "rc_corr.h"
#include "rc_fit_dev.h"
#include "rc_fitvm.h"
#include "rc_fitvm_kern.h"
#include "rc_internal.h"
#include "rcsr_host.h"

#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int CR_TILE = 8192;
constexpr int CR_POLISH = 4;

/* Candidate stages. */
enum {
    CS_NEED_ASM = 1, CS_NEED_SCORE = 2, CS_P_NEED_ASM = 3, CS_P_NEED_SCORE = 4, CS_DONE = 5,
    CS_START_SCORE = 6 /* no iterations: the start's objective only */
};
/* Control operations. */
enum { CO_INIT = 0, CO_AFTER_ASM = 1, CO_AFTER_SCORE = 2, CO_FINAL = 3 };

struct CrState {
    double c[FT_KMAX];
    double trial[FT_KMAX];
    double ajtj[FT_KMAX * (FT_KMAX + 1) / 2]; /* curvature of rho at c */
    double ajtr[FT_KMAX];                     /* gradient of rho at c */
    double rho, best, lambda;
    int stage, it, polish_it, status, sstart, srank, improved, first;
};

__device__ __forceinline__ int tri(int i, int j) { return i * (i + 1) / 2 + j; }

/* Centred weighted second moments of (f, y) from the raw sums:
   c_ff = S_ff - W m_f^2, c_yy = S_yy - W m_y^2, c_fy = S_fy - W m_f m_y. */
struct CrCentred {
    double mf, my, cff, cyy, cfy;
};
__device__ __forceinline__ CrCentred cr_centre(double W, double sf, double sff, double sfy,
                                               double sy, double syy)
{
    CrCentred o;
    o.mf = sf / W;
    o.my = sy / W;
    o.cff = sff - W * o.mf * o.mf;
    o.cyy = syy - W * o.my * o.my;
    o.cfy = sfy - W * o.mf * o.my;
    return o;
}

/* The objective of rho: 1 - rho (signed) or 1 - |rho|. */
__device__ __forceinline__ double cr_objective(double rho, int metric)
{
    return metric == RC_LOSS_CORRELATION_SIGNED ? 1.0 - rho : 1.0 - fabs(rho);
}

/* Score of one constant vector from its reduced moments (a non-finite row
   scores 1e30; no spread in f or y scores 1). */
__device__ __forceinline__ double cr_trial_objective(const double *mo, int bad, int metric)
{
    if (bad) return 1e30;
    const double W = mo[0];
    if (!(W > 0.0)) return 1e30;
    const CrCentred z = cr_centre(W, mo[1], mo[2], mo[3], mo[4], mo[5]);
    if (!(z.cff > 0.0) || !(z.cyy > 0.0)) return 1.0;
    return cr_objective(z.cfy / sqrt(z.cff * z.cyy), metric);
}

/* From the reduced moment Gram G of z = [g | 1 | f | y] (p + 3 columns):
   rho, its gradient in s->ajtr and its Gauss-Newton curvature in s->ajtj,
   the objective and the centred variances.  Without spread the curvature is
   the identity and the gradient zero. */
__device__ void cr_from_gram(CrState *s, const double *G, int p, int metric, double &obj,
                             double &cff, double &cyy)
{
    const int i1 = p, iF = p + 1, iY = p + 2;
    const double W = G[tri(i1, i1)];
    for (int i = 0; i < p; ++i)
        for (int j = 0; j <= i; ++j) s->ajtj[tri(i, j)] = (i == j) ? 1.0 : 0.0;
    for (int j = 0; j < p; ++j) s->ajtr[j] = 0.0;
    double rho = 0.0;
    cff = 0.0;
    cyy = 0.0;
    if (W > 0.0) {
        const CrCentred z = cr_centre(W, G[tri(iF, i1)], G[tri(iF, iF)], G[tri(iY, iF)],
                                      G[tri(iY, i1)], G[tri(iY, iY)]);
        cff = z.cff;
        cyy = z.cyy;
        if (cff > 0.0 && cyy > 0.0) {
            const double root = sqrt(cff * cyy);
            rho = z.cfy / root;
            /* Centred cross moments of each gradient column with f and y. */
            double gm[FT_KMAX], gf[FT_KMAX], gy[FT_KMAX];
            for (int j = 0; j < p; ++j) {
                gm[j] = G[tri(i1, j)] / W;
                gf[j] = G[tri(iF, j)] - W * gm[j] * z.mf;
                gy[j] = G[tri(iY, j)] - W * gm[j] * z.my;
                s->ajtr[j] = gy[j] / root - rho * gf[j] / cff;
            }
            for (int i = 0; i < p; ++i)
                for (int j = 0; j <= i; ++j) {
                    const double cgg = G[tri(i, j)] - W * gm[i] * gm[j];
                    s->ajtj[tri(i, j)] = (cgg - gf[i] * gf[j] / cff) / cff;
                }
        }
    }
    s->rho = rho;
    obj = cr_objective(rho, metric);
}

/* The masked system of one step (a fresh copy of the assembled one): the
   objective's gradient -sigma * d rho / dc and the gauge slots pinned by an
   identity row and column with a zero right-hand side, so their step is 0. */
__device__ void cr_masked_system(const CrState *s, int p, int metric, int gs, int go, double *jtj,
                                 double *jtr)
{
    const double sigma =
        metric == RC_LOSS_CORRELATION_SIGNED ? 1.0 : (s->rho >= 0.0 ? 1.0 : -1.0);
    for (int i = 0; i < p; ++i) {
        const bool pin_i = i == gs || i == go;
        jtr[i] = pin_i ? 0.0 : -sigma * s->ajtr[i];
        for (int j = 0; j <= i; ++j) {
            const bool pin = pin_i || j == gs || j == go;
            jtj[tri(i, j)] = pin ? (i == j ? 1.0 : 0.0) : s->ajtj[tri(i, j)];
        }
    }
}

__device__ __forceinline__ void cr_stage_constants(double *__restrict__ cp, int cand, int p,
                                                   const double *v)
{
    for (int j = 0; j < FT_KMAX; ++j) cp[(size_t)cand * FT_KMAX + j] = j < p ? v[j] : 0.0;
}

/* One wave per candidate; the state is staged in LDS for the launch. */
__global__ __launch_bounds__(64) void cr_ctrl_kernel(
    int op, int ncand, CrState *__restrict__ S, const CrCand *__restrict__ cands,
    const double *__restrict__ c0, int kmax, int iters, int metric, int *__restrict__ run_asm,
    int *__restrict__ run_score, double *__restrict__ cp, const double *__restrict__ pgram,
    const int *__restrict__ pbad_a, const double *__restrict__ pmom,
    const int *__restrict__ pbad_s, double *__restrict__ outc, double *__restrict__ outsse,
    int *__restrict__ outstatus)
{
    const int cand = blockIdx.x;
    if (cand >= ncand) return;
    const int lane = threadIdx.x;
    const CrCand d = cands[cand];
    const int p = d.k;
    __shared__ CrState sS;
    __shared__ double G[CR_TRI];
    __shared__ double Jm[FT_KMAX * (FT_KMAX + 1) / 2], rm[FT_KMAX];
    __shared__ double Lw[FT_KMAX * FT_KMAX], dsc[FT_KMAX], rhs[FT_KMAX], delta[FT_KMAX];
    __shared__ double sh_mom[CR_NMOM];
    __shared__ int sh_bad, sh_solve;
    __shared__ double sh_lambda;
    constexpr int SW = (int)(sizeof(CrState) / sizeof(int));
    static_assert(sizeof(CrState) % sizeof(int) == 0, "CrState is copied as ints");
    for (int i = lane; i < SW; i += 64) ((int *)&sS)[i] = ((const int *)(S + cand))[i];
    __syncthreads();
    CrState *s = &sS;
    const int ran_asm = run_asm[cand];
    const int ran_score = run_score[cand];
    const bool asm_step = op == CO_AFTER_ASM && ran_asm &&
                          (s->stage == CS_NEED_ASM || s->stage == CS_P_NEED_ASM);
    const bool score_step = op == CO_AFTER_SCORE && ran_score &&
                            (s->stage == CS_NEED_SCORE || s->stage == CS_P_NEED_SCORE ||
                             s->stage == CS_START_SCORE);
    if (asm_step) {
        const int q = p + 3;
        for (int t = lane; t < q * (q + 1) / 2; t += 64)
            G[t] = rc_ordered_sum(pgram + (size_t)d.item0 * CR_TRI + t, CR_TRI, d.nitems);
    }
    if (lane == 0) {
        int b = 0;
        if (asm_step) b |= rc_batched_or(pbad_a + d.item0, d.nitems);
        if (score_step) {
            b |= rc_batched_or(pbad_s + d.item0, d.nitems);
            for (int qd = 0; qd < CR_NMOM; ++qd)
                sh_mom[qd] = rc_ordered_sum(pmom + (size_t)d.item0 * CR_NMOM + qd, CR_NMOM, d.nitems);
        }
        sh_bad = b;
    }
    __syncthreads();

    if (lane == 0) {
        int solve = 0; /* 1: LM step at s->lambda, 2: undamped polish step */
        auto lm_end = [&]() {
            if (s->status == 0 && s->best > 0.0) {
                s->polish_it = 0;
                s->stage = CS_P_NEED_ASM;
            } else {
                s->stage = CS_DONE;
            }
        };
        switch (op) {
            case CO_INIT: {
                s->status = 0;
                s->sstart = 0;
                s->srank = 0;
                s->improved = 0;
                s->first = 1;
                s->it = 0;
                s->polish_it = 0;
                s->lambda = 1e-3;
                s->best = 1.0;
                s->rho = 0.0;
                if (d.pre >= 0) {
                    s->stage = CS_DONE;
                    break;
                }
                for (int j = 0; j < FT_KMAX; ++j) s->c[j] = j < p ? c0[(size_t)cand * kmax + j] : 0.0;
                /* Gauge renormalisation: scale -> sign(scale), offset -> 0. */
                if (d.gs >= 0 && d.gs < p && isfinite(s->c[d.gs]) && s->c[d.gs] != 0.0)
                    s->c[d.gs] = s->c[d.gs] > 0.0 ? 1.0 : -1.0;
                if (d.go >= 0 && d.go < p) s->c[d.go] = 0.0;
                cr_stage_constants(cp, cand, p, s->c);
                s->stage = iters > 0 ? CS_NEED_ASM : CS_START_SCORE;
                break;
            }
            case CO_AFTER_ASM: {
                if (!asm_step) break;
                double obj, cff, cyy;
                cr_from_gram(s, G, p, metric, obj, cff, cyy);
                const bool bad = sh_bad || !isfinite(obj) || !isfinite(cff) || !isfinite(cyy);
                if (s->stage == CS_P_NEED_ASM) {
                    if (bad) {
                        s->status = 2;
                        s->stage = CS_DONE;
                    } else {
                        solve = 2;
                    }
                    break;
                }
                if (s->first) {
                    s->first = 0;
                    s->sstart = bad ? 0 : 1;
                    if (bad) {
                        s->status = 2;
                        s->stage = CS_DONE;
                        break;
                    }
                    if (!(cff > 0.0) || !(cyy > 0.0)) {
                        /* No spread: the objective is 1 and nothing moves. */
                        s->best = 1.0;
                        s->stage = CS_DONE;
                        break;
                    }
                    s->best = obj;
                } else if (bad) {
                    s->status = 2;
                    s->stage = CS_DONE;
                    break;
                }
                /* f nearly constant against y: ill-conditioned. */
                if (cff < 1e-24 * cyy) {
                    s->status = 1;
                    s->stage = CS_DONE;
                    break;
                }
                solve = 1;
                break;
            }
            case CO_AFTER_SCORE: {
                if (!score_step) break;
                const double t = cr_trial_objective(sh_mom, sh_bad, metric);
                if (s->stage == CS_START_SCORE) {
                    s->best = t;
                    s->stage = CS_DONE;
                    break;
                }
                if (s->stage == CS_P_NEED_SCORE) {
                    if (t < s->best) {
                        for (int j = 0; j < p; ++j) s->c[j] = s->trial[j];
                        s->best = t;
                        s->improved = 1;
                        s->polish_it += 1;
                        s->stage = (s->polish_it < CR_POLISH && s->best > 0.0) ? CS_P_NEED_ASM
                                                                             : CS_DONE;
                    } else {
                        s->stage = CS_DONE;
                    }
                    if (s->stage == CS_P_NEED_ASM) cr_stage_constants(cp, cand, p, s->c);
                    break;
                }
                s->it += 1;
                bool stop = false, resolve = false;
                if (t < s->best) {
                    for (int j = 0; j < p; ++j) s->c[j] = s->trial[j];
                    s->best = t;
                    s->lambda = fmax(s->lambda / 8.0, 1e-12);
                    s->improved = 1;
                    stop = s->best <= 0.0;
                } else {
                    s->lambda *= 6.0;
                    stop = s->lambda > 1e9;
                    resolve = true;
                }
                if (stop || s->it >= iters) {
                    lm_end();
                } else if (resolve) {
                    solve = 1; /* same assembled system, new lambda */
                } else {
                    s->stage = CS_NEED_ASM;
                }
                if (s->stage == CS_NEED_ASM || s->stage == CS_P_NEED_ASM)
                    cr_stage_constants(cp, cand, p, s->c);
                break;
            }
            case CO_FINAL: {
                int st = s->status;
                if (d.pre >= 0) {
                    st = d.pre;
                } else if (st == 1 || st == 2) {
                    /* A failure after an improvement is 5, before one 6,
                       without a usable start 7. */
                    st = !s->sstart ? 7 : (s->improved ? 5 : 6);
                } else if (st == 0 && s->srank) {
                    st = 4;
                }
                for (int j = 0; j < kmax; ++j)
                    outc[(size_t)cand * kmax + j] =
                        (d.pre < 0 && j < p) ? s->c[j] : c0[(size_t)cand * kmax + j];
                outsse[cand] = d.pre >= 0 ? __builtin_nan("") : s->best;
                outstatus[cand] = st;
                break;
            }
            default: break;
        }
        if (solve) cr_masked_system(s, p, metric, d.gs, d.go, Jm, rm);
        sh_solve = solve;
        sh_lambda = solve == 1 ? s->lambda : 0.0;
    }
    __syncthreads();
    if (sh_solve) {
        const int rc = ft_solve_wave(Jm, rm, sh_lambda, p, delta, Lw, dsc, rhs, lane);
        if (lane == 0) {
            if (rc == 1) {
                if (sh_solve == 1) s->status = 1;
                s->stage = CS_DONE;
            } else {
                if (rc == 4) s->srank = 1;
                for (int j = 0; j < p; ++j) s->trial[j] = fmin(1e6, fmax(-1e6, s->c[j] + delta[j]));
                cr_stage_constants(cp, cand, p, s->trial);
                s->stage = sh_solve == 1 ? CS_NEED_SCORE : CS_P_NEED_SCORE;
            }
        }
    }
    if (lane == 0) {
        const int stg = s->stage;
        run_asm[cand] = stg == CS_NEED_ASM || stg == CS_P_NEED_ASM;
        run_score[cand] = stg == CS_NEED_SCORE || stg == CS_P_NEED_SCORE || stg == CS_START_SCORE;
    }
    __syncthreads();
    for (int i = lane; i < SW; i += 64) ((int *)(S + cand))[i] = ((const int *)&sS)[i];
}

size_t cr_take(size_t &cur, size_t bytes)
{
    const size_t o = (cur + 255) & ~(size_t)255;
    cur = o + bytes;
    return o;
}

} /* namespace */

/* The correlation arm's device region and pinned staging, grown at the high
   water mark and kept with the context. */
struct RcCorrState {
    unsigned char *dev = nullptr, *host = nullptr;
    size_t dev_cap = 0, host_cap = 0;
};

void rc_corr_free(RcCtx *c)
{
    if (!c->corr) return;
    if (c->corr->dev) (void)hipFree(c->corr->dev);
    if (c->corr->host) (void)hipHostFree(c->corr->host);
    delete c->corr;
    c->corr = nullptr;
}

/* The correlation arm of rc_fit_batch. */
int rc_corr_run(RcCtx *c, const uint32_t *prog, const uint32_t *off, const int *pcount, int nprog,
                int kmax, const double *c0, const int *gauge, int iters, int metric, double *outc,
                double *outsse, int *outstatus)
{
    RcRows *r = &c->sets[0];
    if (r->n <= 0 || rc_rows_ensure_f64(c, r) != 0) return -1;
    /* Records the kernels can run without a stray read. */
    for (int i = 0; i < nprog; ++i) {
        if (off[i + 1] < off[i]) return -1;
        const uint32_t words = off[i + 1] - off[i];
        if (words == 0 || words % 4 != 0 || words / 4 > FT_MAX_NODES) return -1;
        if (pcount[i] < 1 || pcount[i] > kmax) continue;
        const uint32_t *nd = prog + off[i];
        for (uint32_t q = 0; q < words / 4; ++q) {
            const uint32_t op = nd[4 * q], k0 = nd[4 * q + 1], k1 = nd[4 * q + 2], slot = nd[4 * q + 3];
            if (op == OP_CONST) {
                if (slot >= (uint32_t)pcount[i]) return -1;
            } else if (op == OP_VAR) {
                if ((int)slot < 0 || (int)slot >= r->v) return -1;
            } else if (k0 >= q || ((op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV ||
                                    op == OP_POW) && k1 >= q)) {
                return -1;
            }
        }
    }
    /* Streams: the f64 fit VM (tape and reverse streams for the 2-row f64
       block, value stream for the 4-row value pass) where the program fits,
       else the records themselves. */
    std::vector<CrCand> cands((size_t)nprog);
    std::vector<std::vector<uint32_t>> vf((size_t)nprog), vr((size_t)nprog), vt((size_t)nprog);
    rc_parallel_for(nprog, 32, [&](int lo, int hi) {
        uint32_t bf[2 * 4 * FT_MAX_NODES], br[2 * 8 * FT_MAX_NODES], bt[2 * 4 * FT_MAX_NODES];
        for (int i = lo; i < hi; ++i) {
            CrCand &d = cands[i];
            memset(&d, 0, sizeof(d));
            d.k = pcount[i];
            d.pre = (pcount[i] < 1 || pcount[i] > kmax) ? 1 : -1;
            d.gs = gauge ? gauge[2 * i] : -1;
            d.go = gauge ? gauge[2 * i + 1] : -1;
            if (d.pre >= 0) continue;
            const int words = (int)(off[i + 1] - off[i]);
            int info[6] = {0, 0, 0, 0, 0, 0}, tinfo[6] = {0, 0, 0, 0, 0, 0};
            const int a = rch_fitvm_compile(prog + off[i], words, d.k, r->v, VM64_XS(2) * 8,
                                            VMC_INPUTS_IN_U, bf, (int)(sizeof(bf) / 4), br,
                                            (int)(sizeof(br) / 4), info);
            const int b = rch_fitvm_compile(prog + off[i], words, d.k, r->v, VM64_XS(2) * 8,
                                            VMC_INPUTS_IN_U | VMC_VALUE_ONLY | VMC_SMALL_U, bt,
                                            (int)(sizeof(bt) / 4), nullptr, 0, tinfo);
            if (a == 0 && b == 0) {
                d.vm = 1;
                vf[i].assign(bf, bf + 2 * info[0]);
                vr[i].assign(br, br + 2 * info[1]);
                vt[i].assign(bt, bt + 2 * tinfo[0]);
            }
        }
    });
    std::vector<uint32_t> code;
    std::vector<FtItem> items;
    enum { NA = 4, NS = 2 };
    std::vector<int> alist[NA], slist[NS];
    const int tile = c->tune_fit_tile > 0 ? c->tune_fit_tile : CR_TILE;
    for (int i = 0; i < nprog; ++i) {
        CrCand &d = cands[i];
        if (d.pre >= 0) continue;
        if (d.vm) {
            if (code.size() & 1) code.push_back(0);
            d.fwd_off = (uint32_t)code.size();
            code.insert(code.end(), vf[i].begin(), vf[i].end());
            d.rev_off = (uint32_t)code.size();
            code.insert(code.end(), vr[i].begin(), vr[i].end());
            d.t_off = (uint32_t)code.size();
            code.insert(code.end(), vt[i].begin(), vt[i].end());
            d.m = (int)vf[i].size() / 2;
            d.nrev = (int)vr[i].size() / 2;
            d.t_m = (int)vt[i].size() / 2;
        } else {
            d.fwd_off = d.rev_off = d.t_off = (uint32_t)code.size();
            code.insert(code.end(), prog + off[i], prog + off[i + 1]);
            d.m = d.t_m = (int)((off[i + 1] - off[i]) / 4);
        }
        const int kc = d.k + 3 <= 16 ? 0 : 1;
        const int acls = (d.vm ? 0 : 2) + kc;
        d.item0 = (int)items.size();
        for (int r0 = 0; r0 < r->n; r0 += tile) {
            FtItem it;
            it.cand = i;
            it.row0 = r0;
            it.rows = std::min(tile, r->n - r0);
            it.p64_blk = 0;
            alist[acls].push_back((int)items.size());
            slist[d.vm ? 0 : 1].push_back((int)items.size());
            items.push_back(it);
        }
        d.nitems = (int)items.size() - d.item0;
    }
    if (code.empty()) code.push_back(0);
    std::vector<int> lists;
    int aofs[NA + 1], sofs[NS + 1];
    for (int q = 0; q < NA; ++q) {
        aofs[q] = (int)lists.size();
        lists.insert(lists.end(), alist[q].begin(), alist[q].end());
    }
    aofs[NA] = (int)lists.size();
    for (int q = 0; q < NS; ++q) {
        sofs[q] = (int)lists.size();
        lists.insert(lists.end(), slist[q].begin(), slist[q].end());
    }
    sofs[NS] = (int)lists.size();
    if (lists.empty()) lists.push_back(0);
    const int nitems = std::max<int>(1, (int)items.size());

    size_t cur = 0;
    const size_t o_code = cr_take(cur, code.size() * 4);
    const size_t o_cand = cr_take(cur, cands.size() * sizeof(CrCand));
    const size_t o_items = cr_take(cur, (size_t)nitems * sizeof(FtItem));
    const size_t o_lists = cr_take(cur, lists.size() * 4);
    const size_t o_c0 = cr_take(cur, (size_t)nprog * kmax * 8);
    const size_t in_end = cr_take(cur, 0);
    const size_t o_outc = cr_take(cur, (size_t)nprog * kmax * 8);
    const size_t o_sse = cr_take(cur, (size_t)nprog * 8);
    const size_t o_st = cr_take(cur, (size_t)nprog * 4);
    const size_t io_end = cr_take(cur, 0);
    const size_t o_state = cr_take(cur, (size_t)nprog * sizeof(CrState));
    const size_t o_run_a = cr_take(cur, (size_t)nprog * 4);
    const size_t o_run_s = cr_take(cur, (size_t)nprog * 4);
    const size_t o_cp = cr_take(cur, (size_t)nprog * FT_KMAX * 8);
    const size_t o_pgram = cr_take(cur, (size_t)nitems * CR_TRI * 8);
    const size_t o_pbad_a = cr_take(cur, (size_t)nitems * 4);
    const size_t o_pmom = cr_take(cur, (size_t)nitems * CR_NMOM * 8);
    const size_t o_pbad_s = cr_take(cur, (size_t)nitems * 4);
    const size_t dev_end = cr_take(cur, 0);

    if (!c->corr) c->corr = new RcCorrState();
    RcCorrState *cs = c->corr;
    if (dev_end > cs->dev_cap) {
        if (cs->dev) rc_dev_free(c, cs->dev);
        cs->dev = nullptr;
        cs->dev_cap = 0;
        const size_t cap = dev_end + dev_end / 4;
        if (rc_dev_alloc(c, (void **)&cs->dev, cap) != 0) return -1;
        cs->dev_cap = cap;
    }
    if (io_end > cs->host_cap) {
        if (cs->host) (void)hipHostFree(cs->host);
        cs->host = nullptr;
        cs->host_cap = 0;
        const size_t cap = io_end + io_end / 4;
        if (hipHostMalloc((void **)&cs->host, cap) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        cs->host_cap = cap;
    }
    unsigned char *H = cs->host, *D = cs->dev;
    memcpy(H + o_code, code.data(), code.size() * 4);
    memcpy(H + o_cand, cands.data(), cands.size() * sizeof(CrCand));
    if (!items.empty()) memcpy(H + o_items, items.data(), items.size() * sizeof(FtItem));
    memcpy(H + o_lists, lists.data(), lists.size() * 4);
    memcpy(H + o_c0, c0, (size_t)nprog * kmax * 8);

    FtRows R;
    R.cols = r->cols;
    R.cols64 = r->cols64;
    R.y64 = r->y64;
    R.w64 = r->w64;
    R.n = r->n;
    R.v = r->v;
    hipStream_t s = c->s_fit;
    const uint32_t *d_code = (const uint32_t *)(D + o_code);
    const CrCand *d_cand = (const CrCand *)(D + o_cand);
    const FtItem *d_items = (const FtItem *)(D + o_items);
    const int *d_lists = (const int *)(D + o_lists);
    const double *d_c0 = (const double *)(D + o_c0);
    double *d_outc = (double *)(D + o_outc);
    double *d_sse = (double *)(D + o_sse);
    int *d_st = (int *)(D + o_st);
    CrState *d_state = (CrState *)(D + o_state);
    int *d_run_a = (int *)(D + o_run_a);
    int *d_run_s = (int *)(D + o_run_s);
    double *d_cp = (double *)(D + o_cp);
    double *d_pgram = (double *)(D + o_pgram);
    int *d_pbad_a = (int *)(D + o_pbad_a);
    double *d_pmom = (double *)(D + o_pmom);
    int *d_pbad_s = (int *)(D + o_pbad_s);

    if (hipMemcpyAsync(D, H, in_end, hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    (void)hipEventRecord(c->t0, s);
    auto ctrl = [&](int op) {
        cr_ctrl_kernel<<<nprog, 64, 0, s>>>(op, nprog, d_state, d_cand, d_c0, kmax, iters, metric,
                                            d_run_a, d_run_s, d_cp, d_pgram, d_pbad_a, d_pmom,
                                            d_pbad_s, d_outc, d_sse, d_st);
    };
    auto asm_pass = [&]() {
        for (int q = 0; q < NA; ++q)
            cr_launch_asm((q & 1) ? 32 : 16, q < 2, aofs[q + 1] - aofs[q], s, d_items,
                          d_lists + aofs[q], d_cand, R, d_code, d_run_a, d_cp, d_pgram, d_pbad_a);
    };
    auto score_pass = [&]() {
        for (int q = 0; q < NS; ++q)
            cr_launch_score(q == 0, sofs[q + 1] - sofs[q], s, d_items, d_lists + sofs[q], d_cand,
                            R, d_code, d_run_s, d_cp, d_pmom, d_pbad_s);
    };
    ctrl(CO_INIT);
    const int slots = iters > 0 ? iters + CR_POLISH : 1;
    for (int slot = 0; slot < slots; ++slot) {
        if (iters > 0) {
            asm_pass();
            ctrl(CO_AFTER_ASM);
        }
        score_pass();
        ctrl(CO_AFTER_SCORE);
    }
    ctrl(CO_FINAL);
    (void)hipEventRecord(c->t1, s);
    if (hipGetLastError() != hipSuccess) return -1;
    if (hipMemcpyAsync(H + in_end, D + in_end, io_end - in_end, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    memcpy(outc, H + o_outc, (size_t)nprog * kmax * 8);
    memcpy(outsse, H + o_sse, (size_t)nprog * 8);
    memcpy(outstatus, H + o_st, (size_t)nprog * 4);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    c->stats.kernel_ms = ms;
    return 0;
}
