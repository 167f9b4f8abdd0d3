/*
 * Host-side program builders exported by librcsrhip.so (pure CPU code).
 *
 * They restate the reference's Rust host layer that sits directly above the
 * device ABI, so a C/C++/Python caller can produce exactly the programs the
 * rc_* entry points consume.  Trees are pre-order node lists: ops[i] is an
 * operator code (device opcodes 0..42, plus 43 asinh, 44 acosh, 45 atanh),
 * vals[i] the constant value (CONST) or column index (VAR).
 */
#ifndef RCSR_HOST_H
#define RCSR_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

/* gpu.rs:198-218 compile_program (lower, peak-stack gate at 56, emit, HALT).
   Returns the word count, or -1 unsupported operator, -2 stack too deep
   (*peak holds the depth), -3 malformed tree or buffer too small. */
int rch_compile_program(const int *ops, const double *vals, int nn, unsigned int *out, int cap,
                        int *peak);

/* fit.rs:985-1056 gpu_fit_plan_reason: post-order 4-word records and start
   constants.  Returns the constant count, or -1 no constants, -2 more than 20
   constants, -3 more than 96 nodes, -5 unsupported operator, -6 malformed
   tree or buffer too small. */
int rch_fit_plan(const int *ops, const double *vals, int nn, unsigned int *out_words,
                 int cap_words, double *out_c0, int cap_c, int *nodes);

/* No reference counterpart: the compiler of the mixed-precision fit's
   register-slot VM (rc_fitvm.h), exported so its output can be checked on
   the CPU.  rec: words / 4 post-order fit records (the rch_fit_plan form);
   p constants; nvars inputs; col_bytes: bytes between gradient columns of
   the assembly block; flags: 1 inputs in U slots 0..nvars-1, 2 value only
   (no tape, no reverse stream), 4 an 8-slot U file.  fwd / rev receive
   (x, y) word pairs; info[0..5] = forward pairs, reverse pairs, U slots,
   B slots, saved-adjoint slots, 0.  Returns 0, or -1 malformed records,
   -2 unsupported operator, -3 the program does not fit the VM, -4 an
   output buffer is too small. */
int rch_fitvm_compile(const unsigned int *rec, int words, int p, int nvars, int col_bytes,
                      unsigned int flags, unsigned int *fwd, int cap_f, unsigned int *rev,
                      int cap_r, int *info);

/* No reference counterpart: the evaluator's program translation (what
   rc_set_programs uploads), exported so it can be checked on the CPU.
   prog / off / nprog: the rc_set_programs batch; nvars: resident columns.
   code receives [per-program streams | hoist programs], meta receives
   [stream offset per program | program list grouped by stack class | pairs
   (hoist program offset, constant word to fill)].  derived (may be NULL: no
   derived columns) holds *nderived (unary operator, source column) pairs,
   column nvars + k being pair k; the pairs this batch admits are appended,
   up to cap_derived pairs in all, and *nderived is updated.  info[0..6] =
   code words, meta words, the four class starts and the batch's peak stack
   depth (filled even when a buffer is too small).  Returns 0, -1 bad
   arguments, -2 a buffer is too small. */
int rch_eval_translate(const unsigned int *prog, const unsigned int *off, int nprog, int nvars,
                       unsigned int *code, int cap_code, unsigned int *meta, int cap_meta,
                       unsigned int *derived, int cap_derived, int *nderived, int *info);

/* No reference counterpart: the replay key of a recorded fit graph for a
   region of the given sizes (code words, candidate records, row-set records,
   work items, class-list words, candidates, kmax, Gram triangle stride),
   with a null region base, 30 LM slots and empty class lists.  The key holds
   every region offset, so two batches whose regions end at the same byte
   but place a table elsewhere never replay each other's graph.  Exported so
   the key can be checked on the CPU.  Writes up to cap words to key.
   Returns the word count, -1 bad arguments, -2 cap too small. */
int rch_fit_graph_key(int code_words, int ncands, int nrowsets, int nitems, int nlist_words,
                      int nprog, int kmax, int tri_stride, unsigned long long *key, int cap);

#ifdef __cplusplus
}
#endif
#endif
