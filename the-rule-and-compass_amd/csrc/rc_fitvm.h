/*
 * Instruction encoding of the fit VM: the register-slot form in which the
 * mixed-precision fit executes a candidate's expression (host compiler in
 * rc_fitvm.cpp, device interpreter in rc_fitvm_dev.h).  Plain C so the host
 * compiler, the device code and the CPU tests share one definition.
 *
 * The VM evaluates the reference's post-order fit records (fit.rs:1101-1196,
 * consumed by ft_forward / ft_backward eval.cu:1173-1317) with the same f32
 * operations per node, re-encoded for a wave-uniform interpreter:
 *   - values live in per-lane register files indexed by wave-uniform slots:
 *       U  16 slots: spilled stack operands and the first tape of partials
 *       B  16 slots: the second tape of partials
 *       A   8 slots: saved adjoints of the reverse sweep
 *     slot 15 of U and B and slot 7 of A are write sinks that are never read;
 *   - the host assigns every slot, so the device does no stack arithmetic;
 *   - common node groups are fused into one instruction: x*c, f(.)*c, . * c,
 *     . + c, . * x (commutative operands are ordered so the leaf is on the
 *     right; a*b == b*a and a+b == b+a exactly in IEEE arithmetic, so values
 *     are unchanged);
 *   - a leaf constant's gradient is written by the instruction that consumes
 *     it, straight into the Gram block column of that constant.
 *
 * An instruction is two 32-bit words (x, y).  x holds one-hot handler bits
 * (the interpreter tests them in a fixed frequency order), option bits and,
 * in the reverse stream, the byte offset of a gradient column; y holds slot
 * fields.
 */
#ifndef RC_FITVM_H
#define RC_FITVM_H

/* File sizes and sinks. */
#define VM_U_SLOTS 16
#define VM_B_SLOTS 16
#define VM_A_SLOTS 8
#define VM_U_SINK 15
#define VM_B_SINK 15
#define VM_A_SINK 7
#define VM_MAX_VARS 8

/* Compiler options (rch_fitvm_compile flags). */
#define VMC_INPUTS_IN_U 1u /* inputs live in U slots 0..nvars-1 (no X file)  */
#define VMC_VALUE_ONLY 2u  /* forward only: no tape, no reverse stream       */
#define VMC_SMALL_U 4u     /* an 8-slot U file: slot 7 is the sink           */
#define VM_U_SINK_SMALL 7

/* ---- forward handlers (x bits 0..9) ------------------------------------ */
/* F_UNMC  f = un(tos); U[us] = f; [NEED] B[bs] = f'(tos, f); tos = f * c   */
#define VMF_UNMC (1u << 0)
/* F_VMC   U[us] = tos; tos = X[xv] * c                                     */
#define VMF_VMC (1u << 1)
/* F_ADD   tos = U[ur] + tos                                                 */
#define VMF_ADD (1u << 2)
/* F_MULC  U[us] = tos; tos = tos * c                                        */
#define VMF_MULC (1u << 3)
/* F_PUSHV U[us] = tos; tos = X[xv]                                          */
#define VMF_PUSHV (1u << 4)
/* F_UN    f = un(tos); [NEED] U[us] = f'(tos, f); tos = f                   */
#define VMF_UN (1u << 5)
/* F_MULV  tos = tos * X[xv]                                                 */
#define VMF_MULV (1u << 6)
/* F_ADDC  tos = tos + c                                                     */
#define VMF_ADDC (1u << 7)
/* F_PUSHC U[us] = tos; tos = c                                              */
#define VMF_PUSHC (1u << 8)
/* F_BIN   generic binary: l = U[ur] (or tos with RC / RV), r = tos (or c /
           X[xv]); [D0] U[us] = dl; [D1] B[bs] = dr; tos = l op r            */
#define VMF_BIN (1u << 9)

/* forward options */
#define VMF_NEED (1u << 12)  /* UN / UNMC: the partial is needed             */
#define VMF_SIN (1u << 13)   /* UN / UNMC: the operator is sin               */
#define VMF_COS (1u << 14)   /* ... cos                                      */
#define VMF_EXP (1u << 15)   /* ... exp                                      */
#define VMF_RC (1u << 16)    /* BIN: right operand is the constant c         */
#define VMF_RV (1u << 17)    /* BIN: right operand is the input X[xv]        */
#define VMF_D0 (1u << 18)    /* BIN: store the left partial in U[us]         */
#define VMF_D1 (1u << 19)    /* BIN: store the right partial in B[bs]        */
/* After the handler and the word's U write: tos = U[ur] + tos, the ADD of a
   spilled left operand and the right operand this word finished (on any word
   whose handler does not read ur itself; when the word spills the left
   operand itself, us == ur). */
#define VMF_ACC (1u << 20)

/* ---- reverse handlers (x bits 0..10) ----------------------------------- */
/* R_UNMC_N  g = A*U[ut]; A = (A*c)*B[bt]                                    */
#define VMR_UNMC_N (1u << 0)
/* R_VMC     g = A*X[xv]; A = A[ar]                                          */
#define VMR_VMC (1u << 1)
/* R_PUSH    A[as] = A                                                       */
#define VMR_PUSH (1u << 2)
/* R_MULC_N  g = A*U[ut]; A = A*c                                            */
#define VMR_MULC_N (1u << 3)
/* R_UNMC_P  g = A*U[ut]; A = A[ar]                                          */
#define VMR_UNMC_P (1u << 4)
/* R_UN      A = A*U[ut]                                                     */
#define VMR_UN (1u << 5)
/* R_MULC_P  g = A*U[ut]; A = A[ar]                                          */
#define VMR_MULC_P (1u << 6)
/* R_MULV    A = A*X[xv]                                                     */
#define VMR_MULV (1u << 7)
/* R_ADDC_N  g = A                                                           */
#define VMR_ADDC_N (1u << 8)
/* R_ADDC_P  g = A; A = A[ar]                                                */
#define VMR_ADDC_P (1u << 9)
/* R_GEN     generic unary / binary node (see the option bits)               */
#define VMR_GEN (1u << 10)

/* R_GEN options: child kinds and the pop */
#define VMR_LNODE (1u << 11)  /* left (or only) child continues the sweep   */
#define VMR_LCONST (1u << 12) /* left (or only) child is a constant leaf    */
#define VMR_RNODE (1u << 13)
#define VMR_RCONST (1u << 14)
#define VMR_UNARY (1u << 15)  /* one child: partial U[ut]                   */
#define VMR_POP (1u << 16)    /* no child continues: A = A[ar]              */
/* Every zero-adjoint product is forced to 0 (ft_backward skips a node whose
   adjoint is 0, eval.cu:1248). */

/* The x word of reverse instructions carries the LDS byte offset of the
   gradient column written by the instruction in bits 17..31. */
#define VMR_GOFS_SHIFT 17

/* ---- y fields ----------------------------------------------------------- */
/* forward:  bits 0..7 c byte offset (slot * 4) | 8..11 us | 12..15 ur |
             16..19 bs | 20..22 xv | 23..27 operator code                  */
/* reverse:  bits 0..7 c byte offset | 8..11 ut | 12..15 bt | 16..18 as |
             19..21 ar | 22..24 xv | 25..29 operator code (R_GEN)          */
#define VM_Y_COFS(y) ((y) & 0xffu)
#define VMF_Y_US(y) (((y) >> 8) & 0xfu)
#define VMF_Y_UR(y) (((y) >> 12) & 0xfu)
#define VMF_Y_BS(y) (((y) >> 16) & 0xfu)
#define VMF_Y_XV(y) (((y) >> 20) & 0x7u)
#define VMF_Y_OP(y) (((y) >> 23) & 0x1fu)
#define VMR_Y_UT(y) (((y) >> 8) & 0xfu)
#define VMR_Y_BT(y) (((y) >> 12) & 0xfu)
#define VMR_Y_AS(y) (((y) >> 16) & 0x7u)
#define VMR_Y_AR(y) (((y) >> 19) & 0x7u)
#define VMR_Y_XV(y) (((y) >> 22) & 0x7u)
#define VMR_Y_OP(y) (((y) >> 25) & 0x1fu)

/* R_GEN occupies two word pairs: the first carries the left (or only)
   child's gradient column in x, the second carries the right child's column
   in its x (same shift) and 0 in y. */

#endif
