/*
 * Shared definitions of the correlation fit arm (rc_fit_batch metrics 2 and
 * 3): the candidate and work-item records the host driver and control
 * kernel (rc_corr.hip) hand to the row kernels (rc_corr_kern.hip).
 */
#ifndef RC_CORR_H
#define RC_CORR_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rc_fit_dev.h"

/* Columns of the moment block z = [g_1 .. g_p, 1, f, y]. */
#define CR_ZMAX (FT_KMAX + 3)
#define CR_TRI (CR_ZMAX * (CR_ZMAX + 1) / 2)
/* Per-item moments of one score pass: W, sum w f, sum w f^2, sum w f y,
   sum w y, sum w y^2. */
#define CR_NMOM 6

struct CrCand {
    uint32_t fwd_off; /* tape stream (fit VM) or post-order records */
    uint32_t rev_off; /* reverse stream (fit VM) */
    uint32_t t_off;   /* value stream (fit VM) or post-order records */
    int m;            /* forward word pairs / records */
    int nrev;         /* reverse word pairs */
    int t_m;          /* value-stream word pairs / records */
    int k;            /* constants */
    int vm;           /* 1: the fit VM streams, 0: the records */
    int item0, nitems;
    int pre;          /* >= 0: a status decided on the host */
    int gs, go;       /* gauge slots (scale, offset), -1 for none */
};

/* Moment-block assembly over the items of one list (q = k + 3 columns, kc 16
   or 32 Gram columns, vm: the fit VM or the record sweeps): per item the
   lower triangle of sum w z z^T (stride CR_TRI) and a bad-row flag. */
void cr_launch_asm(int kc, bool vm, int nl, hipStream_t s, const FtItem *items, const int *list,
                   const CrCand *cands, FtRows rows, const uint32_t *code, const int *run,
                   const double *cp, double *pgram, int *pbad);

/* Score-pass moments over the items of one list: CR_NMOM doubles and a
   bad-row flag per item. */
void cr_launch_score(bool vm, int nl, hipStream_t s, const FtItem *items, const int *list,
                     const CrCand *cands, FtRows rows, const uint32_t *code, const int *run,
                     const double *cp, double *pmom, int *pbad);

#endif
