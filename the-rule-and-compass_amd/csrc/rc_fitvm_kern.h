/*
 * Launchers of the fit VM kernels (rc_fitvm_kern.hip) for the fit driver.
 */
#ifndef RC_FITVM_KERN_H
#define RC_FITVM_KERN_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rc_fit_dev.h"

/* LDS column stride (floats) of the assembly block of R rows per lane: the
   block's 64 R rows + 2, which keeps the MFMA operand reads of a half-wave on
   distinct banks.  A candidate's reverse stream carries gradient-column byte
   offsets built with the stride of the row shape its class runs (LM and
   polish passes alike). */
#define VM_XS(R) (64 * (R) + 2)

/* LDS column stride (doubles) of the plain FP64 arm's assembly block of R
   f64 rows per lane: the block's 64 R rows + 4. */
#define VM64_XS(R) (64 * (R) + 4)

/* Assembly over the items of one class list: kc 16 / 32 Gram columns,
   r = 2 or 4 rows per lane, polish = the f64-residual polish.  With r = 4 the
   polish reads each row's f64 prediction from p64 when it is non-null (the
   exact pass at the same constants wrote it there) and otherwise computes it
   on the candidate's value-pass stream; with r = 2 it computes it on the
   tape stream. */
void vm_launch_asm(int kc, int r, bool polish, int nl, hipStream_t s, const FtItem *items,
                   const int *list, const FtCand *cands, const FtRows *rows, const uint32_t *code,
                   const int *run, const float *cp32, const double *cp64, double *pgram,
                   int tri_stride, double *ploss, int *pbad, int *pskip, const double *p64);

/* The plain FP64 arm's assembly on the fit VM (2 f64 rows per lane, FP64
   MFMA Gram): kc 16 / 32 Gram columns; polish selects the skip-and-count
   rule for non-finite gradient rows. */
void vm_launch_asm64(int kc, bool polish, int nl, hipStream_t s, const FtItem *items,
                     const int *list, const FtCand *cands, const FtRows *rows,
                     const uint32_t *code, const int *run, const double *cp64, double *pgram,
                     int tri_stride, double *ploss, int *pbad, int *pskip);

/* Trial (f32, 8 rows per lane) or exact (f64, 4 rows per lane) loss.  An
   exact pass with p64 non-null also stores each live row's f64 prediction at
   p64[64 * item.p64_blk + (row - item.row0)]. */
void vm_launch_trial(bool exact, int nl, hipStream_t s, const FtItem *items, const int *list,
                     const FtCand *cands, const FtRows *rows, const uint32_t *code, const int *run,
                     const float *cp32, const double *cp64, double *ploss, double *p64);

#endif
