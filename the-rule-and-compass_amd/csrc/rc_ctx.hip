/*
 * Context, resident data, telemetry and probe entry points of librcsrhip.so.
 *
 * Reference behaviour: eval.cu:3137-3262 (rc_create / rc_destroy),
 * 3469-3510 (arena report / force sync), 3836-3862 (stats), 3882-3995
 * (f64 fit rows, score weights, fit sign), 4863-4893 (probe).
 *
 * MI355X layout: every row set is resident in HBM for the context's life,
 * column-major so a wave's 64 lanes read 256 contiguous bytes per column.
 * Streams are created non-blocking at the default priority.
 */
#include "rc_internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <thread>

double rc_now_ms()
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec / 1e6;
}

void rc_parallel_for(int n, int min_chunk, const std::function<void(int, int)> &body)
{
    if (n <= 0) return;
    const unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<unsigned>(hw ? hw : 1u, 8u);
    nt = std::max(1, std::min(nt, n / std::max(1, min_chunk)));
    if (nt == 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    pool.reserve((size_t)nt - 1);
    for (int t = 1; t < nt; ++t) {
        const int lo = (int)((long long)n * t / nt), hi = (int)((long long)n * (t + 1) / nt);
        try {
            pool.emplace_back(body, lo, hi);
        } catch (...) {
            body(lo, hi); /* no thread available: run the chunk here */
        }
    }
    body(0, (int)((long long)n / nt));
    for (std::thread &th : pool) th.join();
}

int rc_dev_alloc(RcCtx *c, void **p, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (hipMalloc(p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return -1;
    }
    c->alloc_calls++;
    if (c->warm) c->steady_alloc_calls++;
    return 0;
}

void rc_dev_free(RcCtx *c, void *p)
{
    if (!p) return;
    (void)hipFree(p);
    c->alloc_calls++;
    if (c->warm) c->steady_alloc_calls++;
}

/* Grow a device buffer to hold `need` elements; capacity only grows. */
int rc_grow(RcCtx *c, void **p, size_t *cap, size_t need, size_t elem)
{
    if (need <= *cap && *p) return 0;
    size_t want = need + need / 4 + 64;
    void *q = nullptr;
    if (rc_dev_alloc(c, &q, want * elem) != 0) return -1;
    if (*p) rc_dev_free(c, *p);
    *p = q;
    *cap = want;
    return 0;
}

int rc_grow_pinned(RcCtx *c, void **p, size_t *cap, size_t need)
{
    if (need <= *cap && *p) return 0;
    size_t want = need + need / 4 + 256;
    void *q = nullptr;
    if (hipHostMalloc(&q, want, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    if (*p) (void)hipHostFree(*p);
    *p = q;
    *cap = want;
    c->alloc_calls++;
    if (c->warm) c->steady_alloc_calls++;
    return 0;
}

int rc_rows_upload(RcRows *r, const float *cols, const float *y, int n, int v)
{
    r->n = n;
    r->v = v;
    size_t cb = sizeof(float) * (size_t)n * (size_t)(v > 0 ? v : 1);
    size_t yb = sizeof(float) * (size_t)(n > 0 ? n : 1);
    if (hipMalloc((void **)&r->cols, cb) != hipSuccess) { r->cols = nullptr; return -1; }
    if (hipMalloc((void **)&r->y, yb) != hipSuccess) { r->y = nullptr; return -1; }
    if (n > 0 && v > 0 && hipMemcpy(r->cols, cols, sizeof(float) * (size_t)n * v,
                                    hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (n > 0 && hipMemcpy(r->y, y, sizeof(float) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        return -1;
    /* Target moments, as the reference caches them at create time. */
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (double)y[i];
    r->ybar = n > 0 ? s / (double)n : 0.0;
    double ss = 0.0;
    for (int i = 0; i < n; ++i) {
        double d = (double)y[i] - r->ybar;
        ss += d * d;
    }
    r->syy = ss;
    return 0;
}

void rc_rows_free(RcRows *r)
{
    if (r->cols) (void)hipFree(r->cols);
    if (r->y) (void)hipFree(r->y);
    if (r->cols64) (void)hipFree(r->cols64);
    if (r->y64) (void)hipFree(r->y64);
    if (r->w64) (void)hipFree(r->w64);
    *r = RcRows();
}

/* f64 twins cast from the resident f32 rows when the caller never uploaded
   exact rows (eval.cu:1951-1957, 2852-2885). */
__global__ void rc_cast_f32_f64(const float *__restrict__ src, double *__restrict__ dst, long long n)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (double)src[i];
}

int rc_rows_ensure_f64(RcCtx *c, RcRows *r)
{
    if (r->cols64 && r->y64) return 0;
    long long total = (long long)r->n * (long long)r->v;
    if (!r->cols64) {
        if (rc_dev_alloc(c, (void **)&r->cols64, sizeof(double) * (size_t)(total > 0 ? total : 1)) != 0)
            return -1;
        if (total > 0)
            rc_cast_f32_f64<<<(unsigned)((total + 255) / 256), 256, 0, c->s_fit>>>(r->cols, r->cols64, total);
    }
    if (!r->y64) {
        if (rc_dev_alloc(c, (void **)&r->y64, sizeof(double) * (size_t)(r->n > 0 ? r->n : 1)) != 0)
            return -1;
        if (r->n > 0)
            rc_cast_f32_f64<<<(unsigned)((r->n + 255) / 256), 256, 0, c->s_fit>>>(r->y, r->y64, r->n);
    }
    if (hipStreamSynchronize(c->s_fit) != hipSuccess) return -1;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

extern "C" {

void *rc_create(const float *cols, int n, int v, const float *y)
{
    if (n < 0 || v < 0 || (n > 0 && (!cols || !y)) || (n > 0 && v <= 0 && cols == nullptr)) return nullptr;
    RcCtx *c = new (std::nothrow) RcCtx();
    if (!c) return nullptr;
    if (rc_rows_upload(&c->sets[0], cols, y, n, v) != 0) {
        (void)hipGetLastError();
        rc_rows_free(&c->sets[0]);
        delete c;
        return nullptr;
    }
    const char *force_sync = getenv("RCSR_GPU_SYNC");
    if (force_sync && force_sync[0] && strcmp(force_sync, "0") != 0 &&
        strcmp(force_sync, "false") != 0 && strcmp(force_sync, "off") != 0) {
        c->async_copy = 0;
        snprintf(c->sync_reason, sizeof(c->sync_reason),
                 "RCSR_GPU_SYNC requests the synchronous copy path");
    }
    if (hipStreamCreateWithFlags(&c->s_eval, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_fit, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->t0) != hipSuccess || hipEventCreate(&c->t1) != hipSuccess ||
        hipEventCreate(&c->tp) != hipSuccess || hipEventCreate(&c->tq0) != hipSuccess ||
        hipEventCreate(&c->tq1) != hipSuccess || hipEventCreate(&c->tq_prep) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_done, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        rc_rows_free(&c->sets[0]);
        delete c;
        return nullptr;
    }
    /* The device address of rc_eval_static's array on this device (a failure
       here only makes that entry point retry on its first call). */
    (void)rc_static_scratch_init();
    return c;
}

void rc_destroy(void *ctx)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return;
    (void)hipDeviceSynchronize();
    for (int i = 0; i < c->nsets; ++i) rc_rows_free(&c->sets[i]);
    if (c->d_tsign) (void)hipFree(c->d_tsign);
    if (c->d_tsign64) (void)hipFree(c->d_tsign64);
    for (int i = 0; i < 2; ++i) {
        if (c->arena_code[i]) (void)hipFree(c->arena_code[i]);
        if (c->arena_meta[i]) (void)hipFree(c->arena_meta[i]);
        if (c->pool_raw[i]) (void)hipFree(c->pool_raw[i]);
    }
    if (c->d_part) (void)hipFree(c->d_part);
    if (c->d_part_bad) (void)hipFree(c->d_part_bad);
    if (c->d_part2) (void)hipFree(c->d_part2);
    if (c->d_pbar) (void)hipFree(c->d_pbar);
    if (c->d_out) (void)hipFree(c->d_out);
    if (c->d_score) (void)hipFree(c->d_score);
    if (c->d_status) (void)hipFree(c->d_status);
    if (c->d_hash) (void)hipFree(c->d_hash);
    if (c->d_derived) (void)hipFree(c->d_derived);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->h_out) (void)hipHostFree(c->h_out);
    rc_fit_state_free(c);
    rc_corr_free(c);
    rc_vary_free(c);
    rc_order_free(c);
    if (c->s_eval) (void)hipStreamDestroy(c->s_eval);
    if (c->s_fit) (void)hipStreamDestroy(c->s_fit);
    if (c->t0) (void)hipEventDestroy(c->t0);
    if (c->t1) (void)hipEventDestroy(c->t1);
    if (c->tp) (void)hipEventDestroy(c->tp);
    if (c->tq0) (void)hipEventDestroy(c->tq0);
    if (c->tq1) (void)hipEventDestroy(c->tq1);
    if (c->tq_prep) (void)hipEventDestroy(c->tq_prep);
    if (c->ev_done) (void)hipEventDestroy(c->ev_done);
    delete c;
}

int rc_set_fit_data(void *ctx, const double *cols, const double *y, const double *w)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !y || (!cols && c->sets[0].v > 0)) return -1;
    RcRows *r = &c->sets[0];
    rc_order_invalidate(c);
    if (r->cols64) { (void)hipFree(r->cols64); r->cols64 = nullptr; }
    if (r->y64) { (void)hipFree(r->y64); r->y64 = nullptr; }
    if (r->w64) { (void)hipFree(r->w64); r->w64 = nullptr; }
    size_t cb = sizeof(double) * (size_t)r->n * (size_t)r->v;
    size_t yb = sizeof(double) * (size_t)r->n;
    if (rc_dev_alloc(c, (void **)&r->cols64, cb) != 0) return -1;
    if (rc_dev_alloc(c, (void **)&r->y64, yb) != 0) return -1;
    if (cb && hipMemcpy(r->cols64, cols, cb, hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (yb && hipMemcpy(r->y64, y, yb, hipMemcpyHostToDevice) != hipSuccess) return -1;
    if (w) {
        if (rc_dev_alloc(c, (void **)&r->w64, yb) != 0) return -1;
        if (yb && hipMemcpy(r->w64, w, yb, hipMemcpyHostToDevice) != hipSuccess) return -1;
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int rc_set_score_weights(void *ctx, const double *w)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return -1;
    RcRows *r = &c->sets[0];
    if (r->w64) { (void)hipFree(r->w64); r->w64 = nullptr; }
    if (!w) return 0;
    size_t yb = sizeof(double) * (size_t)r->n;
    if (rc_dev_alloc(c, (void **)&r->w64, yb) != 0) return -1;
    if (yb && hipMemcpy(r->w64, w, yb, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(r->w64);
        r->w64 = nullptr;
        return -1;
    }
    return 0;
}

int rc_set_fit_sign(void *ctx, const double *sign)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return -1;
    if (c->d_tsign) { (void)hipFree(c->d_tsign); c->d_tsign = nullptr; }
    if (c->d_tsign64) { (void)hipFree(c->d_tsign64); c->d_tsign64 = nullptr; }
    c->has_tsign = 0;
    if (!sign) return 0;
    const int n = c->sets[0].n;
    std::vector<float> h((size_t)n);
    for (int i = 0; i < n; ++i) h[i] = (float)sign[i];
    if (rc_dev_alloc(c, (void **)&c->d_tsign64, sizeof(double) * (size_t)n) != 0) return -1;
    if (rc_dev_alloc(c, (void **)&c->d_tsign, sizeof(float) * (size_t)n) != 0) return -1;
    if (n && (hipMemcpy(c->d_tsign64, sign, sizeof(double) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(c->d_tsign, h.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess))
        return -1;
    c->has_tsign = 1;
    return 0;
}

int rc_add_rows(void *ctx, const float *cols, int n, int v, const float *y, const double *cols64,
                const double *y64, const double *w)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || n <= 0 || v < 0 || !y || (v > 0 && !cols)) return -1;
    if (c->nsets >= RC_MAX_SETS) return -1;
    RcRows *r = &c->sets[c->nsets];
    if (rc_rows_upload(r, cols, y, n, v) != 0) {
        (void)hipGetLastError();
        rc_rows_free(r);
        return -1;
    }
    size_t cb = sizeof(double) * (size_t)n * (size_t)v;
    size_t yb = sizeof(double) * (size_t)n;
    if (cols64 && y64) {
        if (rc_dev_alloc(c, (void **)&r->cols64, cb) != 0 || rc_dev_alloc(c, (void **)&r->y64, yb) != 0 ||
            (cb && hipMemcpy(r->cols64, cols64, cb, hipMemcpyHostToDevice) != hipSuccess) ||
            hipMemcpy(r->y64, y64, yb, hipMemcpyHostToDevice) != hipSuccess) {
            rc_rows_free(r);
            return -1;
        }
    }
    if (w) {
        if (rc_dev_alloc(c, (void **)&r->w64, yb) != 0 ||
            hipMemcpy(r->w64, w, yb, hipMemcpyHostToDevice) != hipSuccess) {
            rc_rows_free(r);
            return -1;
        }
    }
    return c->nsets++;
}

void rc_stats_begin(void *ctx)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return;
    memset(&c->stats, 0, sizeof(c->stats));
    c->batch_start_ms = rc_now_ms();
}

int rc_last_stats(void *ctx, RcStats *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out) return -1;
    *out = c->stats;
    return 0;
}

int rc_arena_report(void *ctx, RcArenaReport *out)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c || !out) return -1;
    memset(out, 0, sizeof(*out));
    out->async_copy = c->async_copy;
    snprintf(out->reason, sizeof(out->reason), "%s", c->sync_reason);
    out->pinned_bytes = (unsigned long long)(c->h_stage_bytes + c->h_out_bytes);
    out->alloc_calls = c->alloc_calls;
    out->steady_alloc_calls = c->steady_alloc_calls;
    out->overlapped_uploads = c->overlapped_uploads;
    out->batches = c->batches;
    out->fs_alloc_calls = c->fs_alloc_calls;
    out->fs_growths = c->fs_growths;
    out->fs_steady_alloc_calls = c->fs_steady_alloc_calls;
    out->fs_batches = c->fs_batches;
    return 0;
}

int rc_arena_force_sync(void *ctx, int sync)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return -1;
    int previous = c->async_copy;
    if (sync) {
        c->async_copy = 0;
        snprintf(c->sync_reason, sizeof(c->sync_reason), "caller forced the synchronous copy path");
    } else {
        c->async_copy = 1;
        c->sync_reason[0] = '\0';
    }
    return previous;
}

int rc_unified(void) { return 0; }

const char *rc_device_name(void)
{
    static char name[256];
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return "unknown";
    snprintf(name, sizeof(name), "%s (%s, %d CUs, %d MB)", p.name, p.gcnArchName,
             p.multiProcessorCount, (int)(p.totalGlobalMem / (1024 * 1024)));
    return name;
}

int rc_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return -1;
    }
    return n;
}

/* Launch shapes (no counterpart in the reference; the autotuner sets them). */
int rc_set_tuning(void *ctx, int key, int value)
{
    RcCtx *c = (RcCtx *)ctx;
    if (!c) return -1;
    const bool tile_ok = value == 0 || (value >= 256 && value <= (1 << 22) && value % 256 == 0);
    int *slot = nullptr;
    switch (key) {
        case RC_TUNE_FIT_TILE_ROWS:
            if (!tile_ok) return -1;
            slot = &c->tune_fit_tile;
            break;
        case RC_TUNE_FIT_ROWS_PER_LANE:
            if (value != 0 && value != 2 && value != 4) return -1;
            slot = &c->tune_fit_rows;
            break;
        case RC_TUNE_EVAL_TILE_ROWS:
            if (!tile_ok) return -1;
            slot = &c->tune_eval_tile;
            break;
        default: return -1;
    }
    const int prev = *slot;
    *slot = value;
    return prev;
}

int rc_get_tuning(void *ctx, int key)
{
    const RcCtx *c = (const RcCtx *)ctx;
    if (!c) return -1;
    switch (key) {
        case RC_TUNE_FIT_TILE_ROWS: return c->tune_fit_tile;
        case RC_TUNE_FIT_ROWS_PER_LANE: return c->tune_fit_rows;
        case RC_TUNE_EVAL_TILE_ROWS: return c->tune_eval_tile;
        default: return -1;
    }
}

} /* extern "C" */
