/*
 * Exact order, rank and maximum statistics of candidate predictions on
 * gfx950: the device stages behind rc_order_stat_score_batch,
 * rc_rank_score_batch and rc_max_scan_batch (reference eval.cu:2888-3133,
 * 4093-4850).
 *
 * Every call scores nsets x ntrials segments: segment seg = set * ntrials +
 * trial evaluates set `set`'s post-order fit records (the rc_fit_batch
 * program form) over the resident f64 rows with the constants c0[seg * kmax
 * ...], in f64.  The statistics are exact functions of the multiset of
 * values (median, maximum) or of integer ranks (rank sums), so they do not
 * depend on how the work is split; the interquartile mean is summed left to
 * right over the sorted window, the CPU's order, so it agrees bit for bit.
 *
 * MI355X layout:
 *  - one thread per (segment, row) evaluates the records into a private value
 *    array and writes one 64-bit sort key (and the row, for ranks);
 *  - rocPRIM's segmented radix sort (through hipCUB) orders each segment;
 *  - order statistics read the sorted keys directly (one wave per segment);
 *  - average ranks come from two scans per segment in one workgroup: a
 *    forward max-scan finds each position's tie-run start, a backward
 *    min-scan its run end, and A = start + end - n is the doubled centred
 *    rank; its products with the target's cached ranks accumulate in exact
 *    128-bit integers (two 64-bit halves), so the reduction order is free;
 *  - the maximum scan is one workgroup per segment reducing an integer key
 *    (|r| bit patterns order like the values), the lowest maximising row and
 *    the feasibility interval of a level along one constant.
 */
#include "rc_fit_dev.h"
#include "rc_internal.h"

#include <hipcub/hipcub.hpp>

#include <string.h>

#include <algorithm>
#include <vector>

namespace {

constexpr int RK_THREADS = 256;
constexpr int MX_THREADS = 256;
constexpr unsigned long long KEY_BAD = 0xFFFFFFFFFFFFFFFFull;

/* Order-preserving 64-bit key of a finite f64 (-0.0 keyed as +0.0, so the two
   form one tie, as f64 == has them). */
__device__ __forceinline__ unsigned long long rk_key(double v)
{
    if (v == 0.0) v = 0.0;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double os_value(unsigned long long k)
{
    return __longlong_as_double((long long)k);
}

/* Prediction of (segment, row): the records of the segment's set with the
   segment's constants. */
__device__ __forceinline__ double seg_predict(const uint32_t *__restrict__ prog,
                                              const uint32_t *__restrict__ off,
                                              const double *__restrict__ c0, int kmax,
                                              const double *__restrict__ cols64, int n, int seg,
                                              int ntrials, int row, double *v)
{
    const int set = seg / ntrials;
    const uint32_t *nd = prog + off[set];
    const int m = (int)((off[set + 1] - off[set]) / 4u);
    ft_rec_values(nd, m, cols64, n, row, c0 + (size_t)seg * (size_t)kmax, v);
    return v[m - 1];
}

/* Sort keys of every (segment, row): |pred - y| bit patterns (order
   statistics, eval.cu:4383-4419) or rank keys of pred with the row as the
   payload (ranks, eval.cu:2952-2982).  A non-finite value keys last and
   flags its segment (status 2). */
template <bool RANK>
__global__ __launch_bounds__(256) void os_keys_kernel(
    const uint32_t *__restrict__ prog, const uint32_t *__restrict__ off,
    const double *__restrict__ c0, int kmax, const double *__restrict__ cols64,
    const double *__restrict__ y64, int n, int nsegs, int ntrials,
    unsigned long long *__restrict__ keys, uint32_t *__restrict__ rows, int *__restrict__ status)
{
    const long long total = (long long)n * (long long)nsegs;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int seg = (int)(i / n);
        const int row = (int)(i - (long long)seg * n);
        double v[FT_MAX_NODES];
        const double pred = seg_predict(prog, off, c0, kmax, cols64, n, seg, ntrials, row, v);
        if (RANK) {
            if (!isfinite(pred)) {
                keys[i] = KEY_BAD;
                status[seg] = 2;
            } else {
                keys[i] = rk_key(pred);
            }
            rows[i] = (uint32_t)row;
        } else {
            const double r = pred - y64[row];
            if (!isfinite(pred) || fabs(pred) > 1e30 || !isfinite(r)) {
                keys[i] = KEY_BAD;
                status[seg] = 2;
            } else {
                keys[i] = (unsigned long long)__double_as_longlong(fabs(r));
            }
        }
    }
}

/* Rank keys of the target (one segment). */
__global__ void rk_target_keys_kernel(const double *__restrict__ y64, int n,
                                      unsigned long long *__restrict__ keys,
                                      uint32_t *__restrict__ rows, int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = y64[i];
    if (!isfinite(v)) {
        keys[i] = KEY_BAD;
        *bad = 1;
    } else {
        keys[i] = rk_key(v);
    }
    rows[i] = (uint32_t)i;
}

/* Median (what 0) or interquartile mean (what 1) of one sorted segment
   (metrics.rs:1193-1223): one wave per segment; the window sum runs left to
   right on one lane, the CPU's order. */
__global__ __launch_bounds__(64) void os_pick_kernel(const unsigned long long *__restrict__ keys,
                                                     int n, int nsegs, int what,
                                                     double *__restrict__ out,
                                                     const int *__restrict__ status)
{
    const int seg = blockIdx.x;
    if (seg >= nsegs || threadIdx.x != 0) return;
    if (status[seg] != 0) {
        out[seg] = __builtin_nan("");
        return;
    }
    const unsigned long long *k = keys + (size_t)seg * (size_t)n;
    if (what == 0) {
        out[seg] = (n & 1) ? os_value(k[n / 2]) : 0.5 * (os_value(k[n / 2 - 1]) + os_value(k[n / 2]));
        return;
    }
    int lo = 0, hi = n;
    if (n >= 4) {
        lo = n / 4;
        hi = max((3 * n) / 4, lo + 1);
    }
    double s = 0.0;
    for (int i = lo; i < hi; ++i) s += os_value(k[i]);
    out[seg] = s / (double)(hi - lo);
}

/* ---- exact 128-bit accumulation ------------------------------------------- */

struct I128 {
    unsigned long long hi, lo;
};

__device__ __forceinline__ void i128_add(I128 &a, const I128 &b)
{
    const unsigned long long lo = a.lo + b.lo;
    a.hi += b.hi + (lo < a.lo ? 1ull : 0ull);
    a.lo = lo;
}

/* Add a signed 64-bit value, sign-extended. */
__device__ __forceinline__ void i128_add64(I128 &a, long long v)
{
    I128 b = {v < 0 ? ~0ull : 0ull, (unsigned long long)v};
    i128_add(a, b);
}

/* Workgroup sum of one I128 per thread (exact, so any order). */
template <int T>
__device__ __forceinline__ I128 i128_block_sum(I128 v, I128 *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (tid < s) {
            I128 a = sh[tid];
            i128_add(a, sh[tid + s]);
            sh[tid] = a;
        }
        __syncthreads();
    }
    const I128 r = sh[0];
    __syncthreads();
    return r;
}

/* ---- average ranks by two scans -------------------------------------------- */

struct MaxCarry {
    int run;
    __device__ int operator()(int block_aggregate)
    {
        const int old = run;
        run = max(run, block_aggregate);
        return old;
    }
};

struct MinCarry {
    int run;
    __device__ int operator()(int block_aggregate)
    {
        const int old = run;
        run = min(run, block_aggregate);
        return old;
    }
};

/* For each sorted segment: pass 1 stores every position's tie-run start
   (the last position <= i whose key differs from its predecessor); pass 2
   walks the segment backwards for the run end and forms the doubled centred
   rank A = start + end + 1 - n (metrics.rs:1258-1286: run [s, e) gets
   s + e - n).  TARGET: write B[row] = A and Sigma B^2; otherwise accumulate
   Sigma A * B[row] and Sigma A^2 into out[4 seg ..] (hi, lo, hi, lo). */
template <bool TARGET>
__global__ __launch_bounds__(RK_THREADS) void rk_runs_kernel(
    const unsigned long long *__restrict__ keys, const uint32_t *__restrict__ rows, int n,
    int nsegs, int *__restrict__ start_of, long long *__restrict__ B,
    unsigned long long *__restrict__ out)
{
    typedef hipcub::BlockScan<int, RK_THREADS> Scan;
    __shared__ typename Scan::TempStorage tmp;
    __shared__ I128 red[RK_THREADS];
    const int seg = blockIdx.x;
    if (seg >= nsegs) return;
    const int tid = threadIdx.x;
    const size_t base = (size_t)seg * (size_t)n;
    const unsigned long long *k = keys + base;
    const uint32_t *rw = rows + base;
    int *st = start_of + base;

    MaxCarry mc = {0};
    for (int c0 = 0; c0 < n; c0 += RK_THREADS) {
        const int i = c0 + tid;
        int f = 0;
        if (i < n) f = (i == 0 || k[i] != k[i - 1]) ? i : 0;
        int s;
        Scan(tmp).InclusiveScan(f, s, hipcub::Max(), mc);
        __syncthreads();
        if (i < n) st[i] = s;
    }
    __syncthreads();

    I128 ab = {0ull, 0ull}, aa = {0ull, 0ull};
    MinCarry nc = {0x7fffffff};
    for (int c0 = ((n - 1) / RK_THREADS) * RK_THREADS; c0 >= 0; c0 -= RK_THREADS) {
        /* Thread t takes position c0 + (T - 1 - t): an inclusive scan over
           threads is then a suffix scan over positions. */
        const int i = c0 + (RK_THREADS - 1 - tid);
        int f = 0x7fffffff;
        if (i < n) f = (i == n - 1 || k[i] != k[i + 1]) ? i : 0x7fffffff;
        int e;
        Scan(tmp).InclusiveScan(f, e, hipcub::Min(), nc);
        __syncthreads();
        if (i < n) {
            const long long A = (long long)st[i] + (long long)e + 1 - (long long)n;
            if (TARGET) {
                B[rw[i]] = A;
                i128_add64(aa, A * A);
            } else {
                i128_add64(ab, A * B[rw[i]]);
                i128_add64(aa, A * A);
            }
        }
    }
    if (!TARGET) ab = i128_block_sum<RK_THREADS>(ab, red);
    aa = i128_block_sum<RK_THREADS>(aa, red);
    if (tid == 0) {
        if (TARGET) {
            out[0] = aa.hi;
            out[1] = aa.lo;
        } else {
            out[4 * (size_t)seg + 0] = ab.hi;
            out[4 * (size_t)seg + 1] = ab.lo;
            out[4 * (size_t)seg + 2] = aa.hi;
            out[4 * (size_t)seg + 3] = aa.lo;
        }
    }
}

/* ---- the exact maximum-error scan ----------------------------------------- */

/* One workgroup per segment (eval.cu:4505-4600): max |r| as an integer
   max of bit patterns with the lowest row on ties; with a constant `coord`
   and a finite level t >= 0, the intersection of the per-row intervals
   {delta : |r_i + b_i delta| <= t}, b_i = d pred_i / d c_coord from the f64
   adjoint sweep (rows with b_i == 0 only veto). */
__global__ __launch_bounds__(MX_THREADS) void mx_scan_kernel(
    const uint32_t *__restrict__ prog, const uint32_t *__restrict__ off,
    const int *__restrict__ pcount, const double *__restrict__ c0, int kmax,
    const int *__restrict__ coords, const double *__restrict__ levels,
    const double *__restrict__ cols64, const double *__restrict__ y64, int n, int nsegs,
    int ntrials, RcMaxOut *__restrict__ out)
{
    __shared__ unsigned long long skey[MX_THREADS];
    __shared__ int srow[MX_THREADS];
    __shared__ double slo[MX_THREADS];
    __shared__ double shi[MX_THREADS];
    __shared__ int sbad[MX_THREADS];
    const int seg = blockIdx.x;
    if (seg >= nsegs) return;
    const int tid = threadIdx.x;
    const int set = seg / ntrials;
    const uint32_t *nd = prog + off[set];
    const int m = (int)((off[set + 1] - off[set]) / 4u);
    const int p = pcount[set];
    const double *c = c0 + (size_t)seg * (size_t)kmax;
    const int coord = (p > 0 && coords[seg] >= 0 && coords[seg] < p) ? coords[seg] : -1;
    const double level = levels[seg];
    const bool want = coord >= 0 && isfinite(level) && level >= 0.0;

    unsigned long long mykey = 0ull;
    int myrow = 0x7fffffff;
    double mylo = -__builtin_inf(), myhi = __builtin_inf();
    int mybad = 0; /* bit 0: a non-finite row; bit 1: the level is vetoed */
    for (int row = tid; row < n; row += MX_THREADS) {
        double v[FT_MAX_NODES];
        ft_rec_values(nd, m, cols64, n, row, c, v);
        const double pred = v[m - 1];
        const double r = pred - y64[row];
        if (!isfinite(pred) || fabs(pred) > 1e30 || !isfinite(r)) {
            mybad |= 1;
            break;
        }
        const unsigned long long key = (unsigned long long)__double_as_longlong(fabs(r));
        if (key > mykey || (key == mykey && row < myrow)) {
            mykey = key;
            myrow = row;
        }
        if (want && !(mybad & 2)) {
            double adj[FT_MAX_NODES];
            double g[FT_KMAX];
            ft_rec_gradient(nd, m, v, adj, g, p);
            const double b = g[coord];
            if (!isfinite(b)) {
                mybad |= 1;
                break;
            }
            if (b == 0.0) {
                if (fabs(r) > level) mybad |= 2;
            } else {
                const double e1 = (-level - r) / b;
                const double e2 = (level - r) / b;
                mylo = fmax(mylo, fmin(e1, e2));
                myhi = fmin(myhi, fmax(e1, e2));
            }
        }
    }
    skey[tid] = mykey;
    srow[tid] = myrow;
    slo[tid] = mylo;
    shi[tid] = myhi;
    sbad[tid] = mybad;
    __syncthreads();
    for (int s = MX_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const int j = tid + s;
            if (skey[j] > skey[tid] || (skey[j] == skey[tid] && srow[j] < srow[tid])) {
                skey[tid] = skey[j];
                srow[tid] = srow[j];
            }
            slo[tid] = fmax(slo[tid], slo[j]);
            shi[tid] = fmin(shi[tid], shi[j]);
            sbad[tid] |= sbad[j];
        }
        __syncthreads();
    }
    if (tid == 0) {
        const int bad = sbad[0];
        RcMaxOut o;
        o.status = (bad & 1) ? 2 : 0;
        o.vmax = (bad & 1) ? __builtin_nan("") : os_value(skey[0]);
        o.row = (bad & 1) ? -1 : srow[0];
        o.lo = (bad & 2) ? __builtin_inf() : slo[0];
        o.hi = (bad & 2) ? -__builtin_inf() : shi[0];
        out[seg] = o;
    }
}

/* ---- host side ------------------------------------------------------------- */

/* Device copies of a call's programs and constants. */
struct StageInput {
    uint32_t *prog = nullptr;
    uint32_t *off = nullptr;
    int *pcount = nullptr;
    double *c0 = nullptr;
};

int upload(RcCtx *c, void **dst, size_t *cap, const void *src, size_t bytes)
{
    if (rc_grow(c, dst, cap, bytes > 0 ? bytes : 1, 1) != 0) return -1;
    if (bytes && hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, c->s_fit) != hipSuccess)
        return -1;
    return 0;
}

int stage_inputs(RcCtx *c, const unsigned int *prog, const unsigned int *off, const int *pcount,
                 int nsets, int nsegs, int kmax, const double *c0, StageInput *in)
{
    RcOrderBufs &b = c->order;
    if (upload(c, (void **)&b.prog, &b.prog_cap, prog, sizeof(uint32_t) * (size_t)off[nsets]) != 0 ||
        upload(c, (void **)&b.off, &b.off_cap, off, sizeof(uint32_t) * (size_t)(nsets + 1)) != 0 ||
        upload(c, (void **)&b.pcount, &b.pcount_cap, pcount, sizeof(int) * (size_t)nsets) != 0 ||
        upload(c, (void **)&b.c0, &b.c0_cap, c0, sizeof(double) * (size_t)nsegs * (size_t)kmax) != 0)
        return -1;
    in->prog = b.prog;
    in->off = b.off;
    in->pcount = b.pcount;
    in->c0 = b.c0;
    return 0;
}

/* Segment offsets 0, n, 2n, ... for the segmented sort (the host copy stays
   alive with the context, so the asynchronous upload never reads freed
   memory). */
int segment_offsets(RcCtx *c, int nsegs, int n)
{
    std::vector<int> &h = c->order.h_seg;
    h.resize((size_t)nsegs + 1);
    for (int s = 0; s <= nsegs; ++s) h[s] = s * n;
    return upload(c, (void **)&c->order.seg, &c->order.seg_cap, h.data(), sizeof(int) * h.size());
}

/* Sort buffers for `total` keys (with row payloads when `pairs`). */
int sort_buffers(RcCtx *c, long long total, bool pairs)
{
    RcOrderBufs &b = c->order;
    const size_t t = (size_t)total;
    if (rc_grow(c, (void **)&b.keys, &b.keys_cap, t, sizeof(unsigned long long)) != 0 ||
        rc_grow(c, (void **)&b.keys_out, &b.keys_out_cap, t, sizeof(unsigned long long)) != 0)
        return -1;
    if (pairs &&
        (rc_grow(c, (void **)&b.rows, &b.rows_cap, t, sizeof(uint32_t)) != 0 ||
         rc_grow(c, (void **)&b.rows_out, &b.rows_out_cap, t, sizeof(uint32_t)) != 0 ||
         rc_grow(c, (void **)&b.start_of, &b.start_cap, t, sizeof(int)) != 0))
        return -1;
    return 0;
}

int segmented_sort(RcCtx *c, long long total, int nsegs, bool pairs)
{
    RcOrderBufs &b = c->order;
    size_t need = 0;
    hipError_t e;
    if (pairs)
        e = hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, need, b.keys, b.keys_out, b.rows,
                                                        b.rows_out, (int)total, nsegs, b.seg,
                                                        b.seg + 1, 0, 64, c->s_fit);
    else
        e = hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, need, b.keys, b.keys_out, (int)total,
                                                       nsegs, b.seg, b.seg + 1, 0, 64, c->s_fit);
    if (e != hipSuccess) return -1;
    if (rc_grow(c, (void **)&b.temp, &b.temp_cap, need > 0 ? need : 1, 1) != 0) return -1;
    size_t tb = b.temp_cap;
    if (pairs)
        e = hipcub::DeviceSegmentedRadixSort::SortPairs(b.temp, tb, b.keys, b.keys_out, b.rows,
                                                        b.rows_out, (int)total, nsegs, b.seg,
                                                        b.seg + 1, 0, 64, c->s_fit);
    else
        e = hipcub::DeviceSegmentedRadixSort::SortKeys(b.temp, tb, b.keys, b.keys_out, (int)total,
                                                       nsegs, b.seg, b.seg + 1, 0, 64, c->s_fit);
    return e == hipSuccess ? 0 : -1;
}

unsigned grid_for(long long total)
{
    long long blocks = (total + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    return (unsigned)(blocks < 1 ? 1 : blocks);
}

/* The target's doubled centred ranks B (cached until the f64 rows change)
   and Sigma B^2. */
int rank_prepare(RcCtx *c)
{
    RcOrderBufs &b = c->order;
    if (b.rank_ready) return 0;
    RcRows *r = &c->sets[0];
    const int n = r->n;
    if (n <= 1 || rc_rows_ensure_f64(c, r) != 0) return -1;
    if (sort_buffers(c, n, true) != 0 || segment_offsets(c, 1, n) != 0) return -1;
    if (rc_grow(c, (void **)&b.B, &b.B_cap, (size_t)n, sizeof(long long)) != 0 ||
        rc_grow(c, (void **)&b.flag, &b.flag_cap, 1, sizeof(int)) != 0 ||
        rc_grow(c, (void **)&b.bb, &b.bb_cap, 2, sizeof(unsigned long long)) != 0)
        return -1;
    if (hipMemsetAsync(b.flag, 0, sizeof(int), c->s_fit) != hipSuccess) return -1;
    rk_target_keys_kernel<<<(unsigned)((n + 255) / 256), 256, 0, c->s_fit>>>(r->y64, n, b.keys,
                                                                            b.rows, b.flag);
    if (hipGetLastError() != hipSuccess || segmented_sort(c, n, 1, true) != 0) return -1;
    rk_runs_kernel<true><<<1, RK_THREADS, 0, c->s_fit>>>(b.keys_out, b.rows_out, n, 1, b.start_of,
                                                          b.B, b.bb);
    if (hipGetLastError() != hipSuccess) return -1;
    int bad = 0;
    if (hipMemcpyAsync(&bad, b.flag, sizeof(int), hipMemcpyDeviceToHost, c->s_fit) != hipSuccess ||
        hipMemcpyAsync(b.h_bb, b.bb, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       c->s_fit) != hipSuccess ||
        hipStreamSynchronize(c->s_fit) != hipSuccess)
        return -1;
    if (bad) return -1; /* a non-finite target has no ranks (metrics.rs:1267) */
    b.rank_ready = 1;
    return 0;
}

/* The record operators with two children (ft_is_binary's set). */
bool is_binary_code(uint32_t op)
{
    return op == OP_ADD || op == OP_SUB || op == OP_MUL || op == OP_DIV || op == OP_POW;
}

/* The records must be a post-order program the kernels can run without a
   stray read: children before parents, constant slots below pcount, input
   columns below v, at most 96 nodes. */
bool stage_args_ok(const RcCtx *c, const unsigned int *prog, const unsigned int *off,
                   const int *pcount, int nsets, int ntrials, int kmax, const double *c0)
{
    if (!c || !prog || !off || !pcount || !c0 || nsets <= 0 || ntrials <= 0 || kmax <= 0 ||
        kmax > FT_KMAX)
        return false;
    const int nv = c->sets[0].v;
    for (int s = 0; s < nsets; ++s) {
        if (off[s + 1] < off[s]) return false;
        const uint32_t words = off[s + 1] - off[s];
        if (words == 0 || words % 4 != 0 || words / 4 > FT_MAX_NODES) return false;
        if (pcount[s] < 0 || pcount[s] > kmax) return false;
        const unsigned int *nd = prog + off[s];
        for (uint32_t i = 0; i < words / 4; ++i) {
            const uint32_t op = nd[4 * i], k0 = nd[4 * i + 1], k1 = nd[4 * i + 2],
                           slot = nd[4 * i + 3];
            if (op == OP_CONST) {
                if (slot >= (uint32_t)pcount[s]) return false;
            } else if (op == OP_VAR) {
                if ((int)slot < 0 || (int)slot >= nv) return false;
            } else if (k0 >= i || (is_binary_code(op) && k1 >= i)) {
                return false;
            }
        }
    }
    return true;
}

} /* namespace */

int rc_order_invalidate(RcCtx *c)
{
    c->order.rank_ready = 0;
    return 0;
}

void rc_order_free(RcCtx *c)
{
    RcOrderBufs &b = c->order;
    void *ptrs[] = {b.prog, b.off, b.pcount, b.c0, b.seg, b.keys, b.keys_out, b.rows, b.rows_out,
                    b.start_of, b.temp, b.status, b.val, b.out, b.B, b.flag, b.bb, b.mx, b.coords,
                    b.levels};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    b = RcOrderBufs();
}

extern "C" {

float rc_order_stat_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                                const int *pcount, int nsets, int ntrials, int kmax,
                                const double *c0, int what, double *out_value, int *out_status)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!stage_args_ok(c, prog, off, pcount, nsets, ntrials, kmax, c0) || !out_value ||
        !out_status || (what != 0 && what != 1))
        return -1.0f;
    RcRows *r = &c->sets[0];
    const int n = r->n;
    const int nsegs = nsets * ntrials;
    const long long total = (long long)n * (long long)nsegs;
    if (n <= 0 || (long long)nsets * ntrials > 0x7fffffffLL || total > 0x7fffffffLL) return -1.0f;
    if (rc_rows_ensure_f64(c, r) != 0) return -1.0f;
    RcOrderBufs &b = c->order;
    StageInput in;
    if (stage_inputs(c, prog, off, pcount, nsets, nsegs, kmax, c0, &in) != 0 ||
        segment_offsets(c, nsegs, n) != 0 || sort_buffers(c, total, false) != 0 ||
        rc_grow(c, (void **)&b.status, &b.status_cap, (size_t)nsegs, sizeof(int)) != 0 ||
        rc_grow(c, (void **)&b.val, &b.val_cap, (size_t)nsegs, sizeof(double)) != 0)
        return -1.0f;
    if (hipMemsetAsync(b.status, 0, sizeof(int) * (size_t)nsegs, c->s_fit) != hipSuccess)
        return -1.0f;
    (void)hipEventRecord(c->t0, c->s_fit);
    os_keys_kernel<false><<<grid_for(total), 256, 0, c->s_fit>>>(
        in.prog, in.off, in.c0, kmax, r->cols64, r->y64, n, nsegs, ntrials, b.keys, nullptr,
        b.status);
    if (hipGetLastError() != hipSuccess || segmented_sort(c, total, nsegs, false) != 0)
        return -1.0f;
    os_pick_kernel<<<(unsigned)nsegs, 64, 0, c->s_fit>>>(b.keys_out, n, nsegs, what, b.val,
                                                         b.status);
    if (hipGetLastError() != hipSuccess) return -1.0f;
    (void)hipEventRecord(c->t1, c->s_fit);
    if (hipMemcpyAsync(out_value, b.val, sizeof(double) * (size_t)nsegs, hipMemcpyDeviceToHost,
                       c->s_fit) != hipSuccess ||
        hipMemcpyAsync(out_status, b.status, sizeof(int) * (size_t)nsegs, hipMemcpyDeviceToHost,
                       c->s_fit) != hipSuccess ||
        hipStreamSynchronize(c->s_fit) != hipSuccess)
        return -1.0f;
    float ms = -1.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    return ms;
}

float rc_rank_score_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                          const int *pcount, int nsets, int ntrials, int kmax, const double *c0,
                          unsigned long long *out, int *out_status, unsigned long long *out_bb)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!stage_args_ok(c, prog, off, pcount, nsets, ntrials, kmax, c0) || !out || !out_status ||
        !out_bb)
        return -1.0f;
    RcRows *r = &c->sets[0];
    const int n = r->n;
    const int nsegs = nsets * ntrials;
    const long long total = (long long)n * (long long)nsegs;
    if (n <= 1 || (long long)nsets * ntrials > 0x7fffffffLL || total > 0x7fffffffLL) return -1.0f;
    if (rank_prepare(c) != 0) return -1.0f;
    RcOrderBufs &b = c->order;
    StageInput in;
    if (stage_inputs(c, prog, off, pcount, nsets, nsegs, kmax, c0, &in) != 0 ||
        segment_offsets(c, nsegs, n) != 0 || sort_buffers(c, total, true) != 0 ||
        rc_grow(c, (void **)&b.status, &b.status_cap, (size_t)nsegs, sizeof(int)) != 0 ||
        rc_grow(c, (void **)&b.out, &b.out_cap, 4 * (size_t)nsegs, sizeof(unsigned long long)) != 0)
        return -1.0f;
    if (hipMemsetAsync(b.status, 0, sizeof(int) * (size_t)nsegs, c->s_fit) != hipSuccess)
        return -1.0f;
    (void)hipEventRecord(c->t0, c->s_fit);
    os_keys_kernel<true><<<grid_for(total), 256, 0, c->s_fit>>>(
        in.prog, in.off, in.c0, kmax, r->cols64, r->y64, n, nsegs, ntrials, b.keys, b.rows,
        b.status);
    if (hipGetLastError() != hipSuccess || segmented_sort(c, total, nsegs, true) != 0)
        return -1.0f;
    rk_runs_kernel<false><<<(unsigned)nsegs, RK_THREADS, 0, c->s_fit>>>(
        b.keys_out, b.rows_out, n, nsegs, b.start_of, b.B, b.out);
    if (hipGetLastError() != hipSuccess) return -1.0f;
    (void)hipEventRecord(c->t1, c->s_fit);
    if (hipMemcpyAsync(out, b.out, 4 * sizeof(unsigned long long) * (size_t)nsegs,
                       hipMemcpyDeviceToHost, c->s_fit) != hipSuccess ||
        hipMemcpyAsync(out_status, b.status, sizeof(int) * (size_t)nsegs, hipMemcpyDeviceToHost,
                       c->s_fit) != hipSuccess ||
        hipStreamSynchronize(c->s_fit) != hipSuccess)
        return -1.0f;
    out_bb[0] = b.h_bb[0];
    out_bb[1] = b.h_bb[1];
    float ms = -1.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    return ms;
}

int rc_rank_debug_b(void *ctx, long long *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out || rank_prepare(c) != 0) return -1;
    if (hipMemcpy(out, c->order.B, sizeof(long long) * (size_t)c->sets[0].n,
                  hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    return 0;
}

float rc_max_scan_batch(void *ctx, const unsigned int *prog, const unsigned int *off,
                        const int *pcount, int nsets, int ntrials, int kmax, const double *c0,
                        const int *coords, const double *levels, RcMaxOut *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!stage_args_ok(c, prog, off, pcount, nsets, ntrials, kmax, c0) || !out) return -1.0f;
    RcRows *r = &c->sets[0];
    const int n = r->n;
    const int nsegs = nsets * ntrials;
    if (n <= 0 || (long long)nsets * ntrials > 0x7fffffffLL ||
        (long long)nsegs * (long long)n > 0x7fffffffLL)
        return -1.0f;
    if (rc_rows_ensure_f64(c, r) != 0) return -1.0f;
    RcOrderBufs &b = c->order;
    StageInput in;
    /* Null coords / levels: the maximum only (coordinate -1, level NaN). */
    std::vector<int> &hc = b.h_coords;
    std::vector<double> &hl = b.h_levels;
    hc.assign((size_t)nsegs, -1);
    hl.assign((size_t)nsegs, __builtin_nan(""));
    if (coords && levels) {
        for (int s = 0; s < nsegs; ++s) {
            hc[s] = coords[s];
            hl[s] = levels[s];
        }
    }
    if (stage_inputs(c, prog, off, pcount, nsets, nsegs, kmax, c0, &in) != 0 ||
        upload(c, (void **)&b.coords, &b.coords_cap, hc.data(), sizeof(int) * hc.size()) != 0 ||
        upload(c, (void **)&b.levels, &b.levels_cap, hl.data(), sizeof(double) * hl.size()) != 0 ||
        rc_grow(c, (void **)&b.mx, &b.mx_cap, (size_t)nsegs, sizeof(RcMaxOut)) != 0)
        return -1.0f;
    (void)hipEventRecord(c->t0, c->s_fit);
    mx_scan_kernel<<<(unsigned)nsegs, MX_THREADS, 0, c->s_fit>>>(
        in.prog, in.off, in.pcount, in.c0, kmax, b.coords, b.levels, r->cols64, r->y64, n, nsegs,
        ntrials, b.mx);
    if (hipGetLastError() != hipSuccess) return -1.0f;
    (void)hipEventRecord(c->t1, c->s_fit);
    if (hipMemcpyAsync(out, b.mx, sizeof(RcMaxOut) * (size_t)nsegs, hipMemcpyDeviceToHost,
                       c->s_fit) != hipSuccess ||
        hipStreamSynchronize(c->s_fit) != hipSuccess)
        return -1.0f;
    float ms = -1.0f;
    (void)hipEventElapsedTime(&ms, c->t0, c->t1);
    return ms;
}

} /* extern "C" */
