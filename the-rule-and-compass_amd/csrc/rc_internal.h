/*
 * Host-side context shared by the translation units of librcsrhip.so.
 *
 * One context owns one resident dataset (row set 0, uploaded by rc_create and
 * rc_set_fit_data) plus optional extra row sets for heterogeneous batches, the
 * evaluator's program arena, the fit region and the streams.  It mirrors the
 * role of the reference's RcCtx (eval.cu:138-269) with an MI355X layout:
 * column-major f32 rows for the mixed-precision kernels, column-major f64
 * twins for the exact passes, everything resident in HBM for the context's
 * life.
 */
#ifndef RC_INTERNAL_H
#define RC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <functional>
#include <vector>

#include "../../include/rcsr_hip.h"

#define RC_MAX_STACK 64
#define RC_MAX_SETS 16

/* One resident row set.  cols[var * n + row]. */
struct RcRows {
    int n = 0;
    int v = 0;
    float *cols = nullptr;
    float *y = nullptr;
    double *cols64 = nullptr;
    double *y64 = nullptr;
    double *w64 = nullptr;
    double ybar = 0.0;
    double syy = 0.0;
};

struct RcFitState; /* rc_fit.hip */
struct RcCorrState; /* rc_corr.hip */
struct RcVaryState; /* rc_vary.hip */

/* Buffers of the order / rank / maximum stages (rc_order.hip); device
   pointers grow with rc_grow, the host vectors stage asynchronous uploads. */
struct RcOrderBufs {
    uint32_t *prog = nullptr;   size_t prog_cap = 0;
    uint32_t *off = nullptr;    size_t off_cap = 0;
    int *pcount = nullptr;      size_t pcount_cap = 0;
    double *c0 = nullptr;       size_t c0_cap = 0;
    int *seg = nullptr;         size_t seg_cap = 0;
    unsigned long long *keys = nullptr;     size_t keys_cap = 0;
    unsigned long long *keys_out = nullptr; size_t keys_out_cap = 0;
    uint32_t *rows = nullptr;     size_t rows_cap = 0;
    uint32_t *rows_out = nullptr; size_t rows_out_cap = 0;
    int *start_of = nullptr;    size_t start_cap = 0;
    void *temp = nullptr;       size_t temp_cap = 0;
    int *status = nullptr;      size_t status_cap = 0;
    double *val = nullptr;      size_t val_cap = 0;
    unsigned long long *out = nullptr; size_t out_cap = 0;
    long long *B = nullptr;     size_t B_cap = 0;    /* target's doubled centred ranks */
    int *flag = nullptr;        size_t flag_cap = 0;
    unsigned long long *bb = nullptr; size_t bb_cap = 0;
    unsigned long long h_bb[2] = {0, 0};             /* Sigma B^2 (hi, lo) */
    int rank_ready = 0;
    RcMaxOut *mx = nullptr;     size_t mx_cap = 0;
    int *coords = nullptr;      size_t coords_cap = 0;
    double *levels = nullptr;   size_t levels_cap = 0;
    std::vector<int> h_seg;
    std::vector<int> h_coords;
    std::vector<double> h_levels;
};

struct RcCtx {
    RcRows sets[RC_MAX_SETS];
    int nsets = 1;
    /* Launch shapes set by rc_set_tuning (0 = the built-in choice): rows per
       fit work item, rows per lane of the fit VM assembly, rows per
       evaluation tile. */
    int tune_fit_tile = 0;
    int tune_fit_rows = 0;
    int tune_eval_tile = 0;
    float *d_tsign = nullptr;
    double *d_tsign64 = nullptr;
    int has_tsign = 0;

    /* Evaluator arena: two program slots (translated internal streams) so an
       upload may overlap the previous batch, a persistent output and the
       per-tile partial sums. */
    uint32_t *arena_code[2] = {nullptr, nullptr};
    uint32_t *arena_meta[2] = {nullptr, nullptr}; /* per program: code offset */
    size_t arena_code_words = 0;
    size_t arena_meta_words = 0;
    int arena_slot = 0;
    int nprog = 0;
    int max_stack = 0;
    /* Programs grouped by peak stack depth (<=8, <=16, <=32, <=64): class k
       occupies [cls_start[k], cls_start[k+1]) of the program list stored
       after the code offsets in the meta arena. */
    int cls_start[5] = {0, 0, 0, 0, 0};
    /* Derived columns of the evaluator's rows (rc_eval.hip): pair k =
       (unary operator, source column) is column v + k, stored at
       d_derived + k * n; computed once and kept with the rows. */
    RcOrderBufs order;
    std::vector<uint32_t> derived;
    float *d_derived = nullptr;
    int derived_cols = 0;   /* computed */
    int derived_cap = 0;    /* allocated */
    double *d_part = nullptr; /* per (program, tile) partial sums */
    size_t part_cap = 0;
    int *d_part_bad = nullptr;
    size_t part_bad_cap = 0;
    double *d_part2 = nullptr; /* second per-item sum (weights, cross products) */
    size_t part2_cap = 0;
    double *d_pbar = nullptr;  /* per-program mean prediction (correlation scores) */
    size_t pbar_cap = 0;
    float *d_out = nullptr;
    size_t out_cap = 0;
    double *d_score = nullptr;
    int *d_status = nullptr;
    size_t score_cap = 0;
    uint32_t *h_stage = nullptr; /* pinned upload staging */
    size_t h_stage_bytes = 0;
    float *h_out = nullptr;      /* pinned download staging */
    size_t h_out_bytes = 0;

    /* Fixed-width program pool (rc_pool_*). */
    uint32_t *pool_raw[2] = {nullptr, nullptr};
    int pool_cap = 0;
    int pool_width = 0;
    std::vector<uint32_t> pool_host[2];
    unsigned long long *d_hash = nullptr;

    hipStream_t s_eval = nullptr;
    hipStream_t s_fit = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    /* Recorded after a batch's upload, before its derived columns and hoisted
       constants are computed; the batch's first evaluation is timed from it
       when prep_timed is set. */
    hipEvent_t tp = nullptr;
    int prep_timed = 0;
    /* Events owned by the pending submit/collect batch.  rc_eval_submit
       swaps the upload's prep event into tq_prep, so an upload of the next
       batch while this one runs records into a different event, and a
       synchronous score call in between records its own t0/t1. */
    hipEvent_t tq0 = nullptr, tq1 = nullptr, tq_prep = nullptr;
    int pending_prep = 0;
    hipEvent_t ev_done = nullptr;

    /* Pending evaluator batch (submit/collect split). */
    int pending = 0;
    int pending_loss = 0;
    int pending_nprog = 0;

    RcStats stats{};
    double batch_start_ms = 0.0;

    int async_copy = 1;
    char sync_reason[96] = {0};
    int warm = 0;
    unsigned long long alloc_calls = 0;
    unsigned long long steady_alloc_calls = 0;
    unsigned long long overlapped_uploads = 0;
    unsigned long long batches = 0;
    unsigned long long fs_alloc_calls = 0;
    unsigned long long fs_growths = 0;
    unsigned long long fs_steady_alloc_calls = 0;
    unsigned long long fs_batches = 0;

    RcFitState *fit = nullptr;
    RcCorrState *corr = nullptr; /* the correlation arm's region (rc_corr.hip) */
    RcVaryState *vary = nullptr; /* the device variation's region (rc_vary.hip) */
};

/* Host helpers (rc_ctx.hip). */
double rc_now_ms();
int rc_dev_alloc(RcCtx *c, void **p, size_t bytes);
void rc_dev_free(RcCtx *c, void *p);
int rc_grow(RcCtx *c, void **p, size_t *cap, size_t need, size_t elem);
int rc_rows_ensure_f64(RcCtx *c, RcRows *r);
struct FtRows;
/* rc_fit.hip: the fit kernels' row-set table (resident sets, then the
   logistic arm's view of set 0 with the label signs as its target). */
std::vector<FtRows> rc_fit_rows(RcCtx *c);
/* Run body(lo, hi) over [0, n) in contiguous chunks of at least min_chunk
   indices on up to 8 host threads; the calling thread takes the first
   chunk.  body may write only what belongs to its own indices. */
void rc_parallel_for(int n, int min_chunk, const std::function<void(int, int)> &body);
/* rc_order.hip: drop the cached target ranks (the f64 rows changed) / free
   the stage buffers. */
int rc_order_invalidate(RcCtx *c);
/* rc_corr.hip: the correlation arm of rc_fit_batch (metrics 2 and 3). */
void rc_corr_free(RcCtx *c);
void rc_vary_free(RcCtx *c);
/* rc_eval.hip: look up the current device's copy of rc_eval_static's device
   array (once per device).  Returns 0 or -1. */
int rc_static_scratch_init();
int rc_corr_run(RcCtx *c, const uint32_t *prog, const uint32_t *off, const int *pcount, int nprog,
                int kmax, const double *c0, const int *gauge, int iters, int metric, double *outc,
                double *outsse, int *outstatus);
void rc_order_free(RcCtx *c);
int rc_grow_pinned(RcCtx *c, void **p, size_t *cap, size_t need);
int rc_rows_upload(RcRows *r, const float *cols, const float *y, int n, int v);
void rc_rows_free(RcRows *r);
void rc_fit_state_free(RcCtx *c);

/* Translate the ABI's postfix programs into the evaluator's internal stream
   (rc_eval.hip).  Returns the batch's peak stack depth. */
int rc_translate_programs(const uint32_t *prog, const uint32_t *off, int nprog, int nvars,
                          std::vector<uint32_t> &code, std::vector<uint32_t> &meta,
                          int cls_start[5], std::vector<uint32_t> *derived, int max_derived);

#define RC_CHECK(x) do { if ((x) != hipSuccess) return -1; } while (0)

#endif
