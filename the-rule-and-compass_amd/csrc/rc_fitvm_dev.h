/*
 * Device interpreter of the fit VM (encoding in rc_fitvm.h, compiler in
 * rc_fitvm.cpp), included by rc_fit.hip.
 *
 * One lane carries several rows at once: a "pack" value holds 2 f32 rows
 * (vm_f2), 4 f32 rows (vm_f4) or 2 f64 rows (vm_d2).  The register files
 * store a slot as 64-bit elements of an LLVM vector, so one wave-uniform
 * indexed move (s_set_gpr_idx) reads or writes the slot for two f32 rows at
 * once, and the scalar cost of decoding and dispatching an instruction is
 * shared by every row of the pack.
 *
 * Per node the arithmetic is ft_forward's (eval.cu:1173-1233) for f32 rows
 * and value_of's (fit.rs:1211-1365) for f64 rows; the adjoint products are
 * ft_backward's (eval.cu:1236-1317) with the tape partials of the general
 * interpreter, and a zero adjoint contributes exactly 0.
 */
#ifndef RC_FITVM_DEV_H
#define RC_FITVM_DEV_H

#include "rc_fit_dev.h"
#include "rc_fitvm.h"

typedef unsigned long long vm_u64;
typedef vm_u64 vm_u64x2 __attribute__((ext_vector_type(2)));
typedef vm_u64 vm_u64x8 __attribute__((ext_vector_type(8)));
typedef vm_u64 vm_u64x16 __attribute__((ext_vector_type(16)));
typedef float vm_f2 __attribute__((ext_vector_type(2)));
typedef float vm_f4 __attribute__((ext_vector_type(4)));
typedef double vm_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ vm_f2 vm_unpack(vm_u64 v)
{
    vm_f2 r;
    r.x = __uint_as_float((unsigned)v);
    r.y = __uint_as_float((unsigned)(v >> 32));
    return r;
}

__device__ __forceinline__ vm_u64 vm_pack(vm_f2 x)
{
    return (vm_u64)__float_as_uint(x.x) | ((vm_u64)__float_as_uint(x.y) << 32);
}

/* Register files are plain local vectors of 64-bit elements (a struct or an
   array holding them would be demoted to scratch): a file is up to four such
   vectors (f0..f3) indexed by one slot number.  A vm_f2 pack keeps rows 0-1
   in f0, vm_f4 rows 0-1 / 2-3 in f0 / f1, vm_f8 four row pairs in f0..f3;
   a vm_d2 pack keeps one f64 row in each of f0, f1, vm_d4 one in each of
   f0..f3.  All halves of a slot share one index, so the compiler moves them
   under one s_set_gpr_idx bracket.  The slot index is masked to the file
   size, so an 8-slot file maps the 16-slot sink onto its own sink, slot 7. */
typedef float vm_f8 __attribute__((ext_vector_type(8)));
typedef double vm_d4 __attribute__((ext_vector_type(4)));

template <typename F> struct VmSlots;
template <> struct VmSlots<vm_u64x2> { enum { N = 2 }; };
template <> struct VmSlots<vm_u64x8> { enum { N = 8 }; };
template <> struct VmSlots<vm_u64x16> { enum { N = 16 }; };
template <> struct VmSlots<dv8> { enum { N = 8 }; };
template <> struct VmSlots<dv16> { enum { N = 16 }; };
template <typename F> __device__ __forceinline__ unsigned vm_slot(unsigned i)
{
    return i & (unsigned)(VmSlots<F>::N - 1);
}

template <typename V> struct VmPack;
template <> struct VmPack<vm_f2> {
    typedef float T;
    typedef vm_u64x16 F16;
    typedef vm_u64x8 F8;
    enum { R = 2 };
    template <typename F>
    static __device__ __forceinline__ vm_f2 get(const F &f0, const F &, const F &, const F &, unsigned i)
    {
        return vm_unpack(f0[vm_slot<F>(i)]);
    }
    template <typename F>
    static __device__ __forceinline__ void set(F &f0, F &, F &, F &, unsigned i, vm_f2 x)
    {
        f0[vm_slot<F>(i)] = vm_pack(x);
    }
};
template <> struct VmPack<vm_f4> {
    typedef float T;
    typedef vm_u64x16 F16;
    typedef vm_u64x8 F8;
    enum { R = 4 };
    template <typename F>
    static __device__ __forceinline__ vm_f4 get(const F &f0, const F &f1, const F &, const F &, unsigned i)
    {
        const unsigned j = vm_slot<F>(i);
        const vm_f2 a = vm_unpack(f0[j]), b = vm_unpack(f1[j]);
        return (vm_f4){a.x, a.y, b.x, b.y};
    }
    template <typename F>
    static __device__ __forceinline__ void set(F &f0, F &f1, F &, F &, unsigned i, vm_f4 x)
    {
        const unsigned j = vm_slot<F>(i);
        f0[j] = vm_pack((vm_f2){x[0], x[1]});
        f1[j] = vm_pack((vm_f2){x[2], x[3]});
    }
};
template <> struct VmPack<vm_f8> {
    typedef float T;
    typedef vm_u64x16 F16;
    typedef vm_u64x8 F8;
    enum { R = 8 };
    template <typename F>
    static __device__ __forceinline__ vm_f8 get(const F &f0, const F &f1, const F &f2, const F &f3,
                                                unsigned i)
    {
        const unsigned j = vm_slot<F>(i);
        const vm_f2 a = vm_unpack(f0[j]), b = vm_unpack(f1[j]), c = vm_unpack(f2[j]),
                    d = vm_unpack(f3[j]);
        return (vm_f8){a.x, a.y, b.x, b.y, c.x, c.y, d.x, d.y};
    }
    template <typename F>
    static __device__ __forceinline__ void set(F &f0, F &f1, F &f2, F &f3, unsigned i, vm_f8 x)
    {
        const unsigned j = vm_slot<F>(i);
        f0[j] = vm_pack((vm_f2){x[0], x[1]});
        f1[j] = vm_pack((vm_f2){x[2], x[3]});
        f2[j] = vm_pack((vm_f2){x[4], x[5]});
        f3[j] = vm_pack((vm_f2){x[6], x[7]});
    }
};
template <> struct VmPack<vm_d2> {
    typedef double T;
    typedef dv16 F16;
    typedef dv8 F8;
    enum { R = 2 };
    template <typename F>
    static __device__ __forceinline__ vm_d2 get(const F &f0, const F &f1, const F &, const F &, unsigned i)
    {
        const unsigned j = vm_slot<F>(i);
        return (vm_d2){f0[j], f1[j]};
    }
    template <typename F>
    static __device__ __forceinline__ void set(F &f0, F &f1, F &, F &, unsigned i, vm_d2 x)
    {
        const unsigned j = vm_slot<F>(i);
        f0[j] = x[0];
        f1[j] = x[1];
    }
};
template <> struct VmPack<vm_d4> {
    typedef double T;
    typedef dv16 F16;
    typedef dv8 F8;
    enum { R = 4 };
    template <typename F>
    static __device__ __forceinline__ vm_d4 get(const F &f0, const F &f1, const F &f2, const F &f3,
                                                unsigned i)
    {
        const unsigned j = vm_slot<F>(i);
        return (vm_d4){f0[j], f1[j], f2[j], f3[j]};
    }
    template <typename F>
    static __device__ __forceinline__ void set(F &f0, F &f1, F &f2, F &f3, unsigned i, vm_d4 x)
    {
        const unsigned j = vm_slot<F>(i);
        f0[j] = x[0];
        f1[j] = x[1];
        f2[j] = x[2];
        f3[j] = x[3];
    }
};

template <typename V> struct VmScalar {
    typedef typename VmPack<V>::T T;
    enum { R = VmPack<V>::R };
};

/* Constant c of a y word: the byte offset is slot * 4 (f32 constants). */
__device__ __forceinline__ float vm_const(const float *cst, uint32_t y)
{
    return *(const float *)((const char *)cst + VM_Y_COFS(y));
}
__device__ __forceinline__ double vm_const(const double *cst, uint32_t y)
{
    return *(const double *)((const char *)cst + 2u * VM_Y_COFS(y));
}

/* ---- per-component operator helpers ------------------------------------- */

template <typename V> __device__ __forceinline__ V vm_bcast(typename VmScalar<V>::T c)
{
    return (V)(c);
}

/* The out-of-line operator helpers receive the opcode in a VGPR (the calling
   convention), though it is wave-uniform: reading it back with readfirstlane
   turns their operator dispatch into scalar branches. */
__device__ __forceinline__ unsigned vm_uniform(unsigned op)
{
    return (unsigned)__builtin_amdgcn_readfirstlane((int)op);
}

/* ft_forward / value_of unary semantics on every row of a pack. */
template <typename V> __device__ __noinline__ V vm_unary_call(unsigned op_, V a)
{
    const unsigned op = vm_uniform(op_);
    V r;
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = ft_unary(op, a[q]);
    return r;
}

template <typename V> __device__ __noinline__ V vm_partial_call(unsigned op_, V a, V f)
{
    const unsigned op = vm_uniform(op_);
    V r;
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = ft_unary_partial<typename VmScalar<V>::T>(op, a[q], f[q]);
    return r;
}

template <typename V> __device__ __noinline__ V vm_binary_call(unsigned op_, V a, V b)
{
    const unsigned op = vm_uniform(op_);
    V r;
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = ft_binary(op, a[q], b[q]);
    return r;
}

/* Local partials of a binary node (ft_forward's tape values); MUL's are its
   operands and stay inline, the others are out of line. */
template <typename V> __device__ __noinline__ void vm_binary_partials_call(unsigned op_, V a, V b,
                                                                        V &d0, V &d1)
{
    typedef typename VmScalar<V>::T T;
    const unsigned op = vm_uniform(op_);
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) {
        T p0 = (T)0, p1 = (T)0;
        if (op == OP_DIV) { p0 = (T)1 / b[q]; p1 = -a[q] / (b[q] * b[q]); }
        else if (op == OP_POW && a[q] > (T)0) {
            p0 = b[q] * pow(a[q], b[q] - (T)1);
            p1 = pow(a[q], b[q]) * log(a[q]);
        }
        d0[q] = p0;
        d1[q] = p1;
    }
}

template <typename V> __device__ __forceinline__ void vm_binary_partials(unsigned op, V a, V b, V &d0,
                                                                      V &d1)
{
    if (op == OP_MUL) {
        d0 = b;
        d1 = a;
    } else {
        vm_binary_partials_call<V>(op, a, b, d0, d1);
    }
}

__device__ __forceinline__ float vm_exp1(float a) { return (a > 700.0f) ? __builtin_inff() : expf(a); }
__device__ __forceinline__ double vm_exp1(double a) { return (a > 700.0) ? __builtin_inf() : exp(a); }

/* Unary value with the common operators inline. */
template <typename V> __device__ __forceinline__ V vm_unary(uint32_t x, uint32_t y, V a)
{
    V r;
    if (x & VMF_SIN) {
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = sin(a[q]);
    } else if (x & VMF_COS) {
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = cos(a[q]);
    } else if (x & VMF_EXP) {
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = vm_exp1(a[q]);
    } else {
        r = vm_unary_call<V>(VMF_Y_OP(y), a);
    }
    return r;
}

template <typename V> __device__ __forceinline__ V vm_partial(uint32_t x, uint32_t y, V a, V f)
{
    V r;
    if (x & VMF_SIN) {
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = cos(a[q]);
    } else if (x & VMF_COS) {
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) r[q] = -sin(a[q]);
    } else if (x & VMF_EXP) {
        r = f;
    } else {
        r = vm_partial_call<V>(VMF_Y_OP(y), a, f);
    }
    return r;
}

/* Value and partial of a unary node at once: the sin / cos pair comes from
   one sincos evaluation (a shared argument reduction).  On gfx950 sincosf's
   and sincos's two outputs are the same bits as sinf / cosf and sin / cos
   (checked over 2^24 f32 arguments, |x| < 20, |x| < 2e4 and arbitrary bit
   patterns, and their f64 widenings), so a fit does not change. */
__device__ __forceinline__ void vm_sincos1(float a, float *s, float *c) { sincosf(a, s, c); }
__device__ __forceinline__ void vm_sincos1(double a, double *s, double *c) { sincos(a, s, c); }
template <typename V> __device__ __forceinline__ void vm_unary_and_partial(uint32_t x, uint32_t y, V a,
                                                                          V &f, V &d)
{
    if (x & (VMF_SIN | VMF_COS)) {
        V sn, cs;
#pragma unroll
        for (int q = 0; q < VmScalar<V>::R; ++q) {
            typename VmScalar<V>::T s1, c1;
            vm_sincos1(a[q], &s1, &c1);
            sn[q] = s1;
            cs[q] = c1;
        }
        if (x & VMF_SIN) {
            f = sn;
            d = cs;
        } else {
            f = cs;
            d = -sn;
        }
        return;
    }
    f = vm_unary<V>(x, y, a);
    d = vm_partial<V>(x, y, a, f);
}

/* One instruction word of a stream at byte offset o (a wave-uniform scalar
   load). */
__device__ __forceinline__ uint2 vm_word(const uint2 *__restrict__ s, uint32_t o)
{
    return *(const uint2 *)((const char *)s + o);
}

/* ---- forward sweep ----------------------------------------------------- */

/* Runs the forward stream on one pack of rows.  With TAPE the partials the
   reverse stream reads are left in U and B.  Every step writes U once (the
   sink slot when it has nothing to keep), so no branch join merges two
   versions of the file. */
template <typename V, bool TAPE, typename FX, typename FU, typename FB>
__device__ __forceinline__ V vm_forward(const uint2 *__restrict__ fc, int nf,
                                        const typename VmScalar<V>::T *__restrict__ cst,
                                        const FX &X0, const FX &X1, const FX &X2, const FX &X3,
                                        FU &U0, FU &U1, FU &U2, FU &U3, FB &B0, FB &B1, FB &B2,
                                        FB &B3)
{
    typedef VmPack<V> P;
    V tos = vm_bcast<V>(0);
    const uint32_t end = 8u * (uint32_t)nf;
    for (uint32_t o = 0; o < end; o += 8u) {
        const uint2 w = vm_word(fc, o);
        const uint32_t x = w.x, y = w.y;
        V nt, uv = tos;
        uint32_t us = VMF_Y_US(y);
        /* Exactly one handler bit per word: the tests run in the order of
           the handlers' frequency in fitted programs. */
        if (x & VMF_VMC) {
            nt = P::get(X0, X1, X2, X3, VMF_Y_XV(y)) * vm_const(cst, y);
        } else if (x & VMF_UNMC) {
            V f;
            if (TAPE && (x & VMF_NEED)) {
                V dv;
                vm_unary_and_partial<V>(x, y, tos, f, dv);
                P::set(B0, B1, B2, B3, VMF_Y_BS(y), dv);
            } else {
                f = vm_unary<V>(x, y, tos);
            }
            uv = f;
            nt = f * vm_const(cst, y);
        } else if (x & VMF_MULV) {
            nt = tos * P::get(X0, X1, X2, X3, VMF_Y_XV(y));
        } else if (x & VMF_PUSHV) {
            nt = P::get(X0, X1, X2, X3, VMF_Y_XV(y));
        } else if (x & VMF_ADDC) {
            nt = tos + vm_const(cst, y);
        } else if (x & VMF_MULC) {
            nt = tos * vm_const(cst, y);
        } else if (x & VMF_UN) {
            V f;
            if (TAPE && (x & VMF_NEED)) vm_unary_and_partial<V>(x, y, tos, f, uv);
            else f = vm_unary<V>(x, y, tos);
            nt = f;
        } else if (x & VMF_ADD) {
            nt = P::get(U0, U1, U2, U3, VMF_Y_UR(y)) + tos;
        } else if (x & VMF_PUSHC) {
            nt = vm_bcast<V>(vm_const(cst, y));
        } else {
            V l, r;
            if (x & VMF_RC) {
                l = tos;
                r = vm_bcast<V>(vm_const(cst, y));
            } else if (x & VMF_RV) {
                l = tos;
                r = P::get(X0, X1, X2, X3, VMF_Y_XV(y));
            } else {
                l = P::get(U0, U1, U2, U3, VMF_Y_UR(y));
                r = tos;
            }
            const unsigned op = VMF_Y_OP(y);
            if (op == OP_ADD) nt = l + r;
            else if (op == OP_SUB) nt = l - r;
            else if (op == OP_MUL) nt = l * r;
            else nt = vm_binary_call<V>(op, l, r);
            if (TAPE && (x & (VMF_D0 | VMF_D1))) {
                V d0, d1;
                vm_binary_partials<V>(op, l, r, d0, d1);
                uv = d0;
                if (x & VMF_D1) P::set(B0, B1, B2, B3, VMF_Y_BS(y), d1);
            }
        }
        P::set(U0, U1, U2, U3, us, uv);
        /* After the U write: a word that spills the left operand itself
           (us == ur) has just stored it. */
        if (x & VMF_ACC) nt = P::get(U0, U1, U2, U3, VMF_Y_UR(y)) + nt;
        tos = nt;
    }
    return tos;
}

/* ---- reverse sweep ----------------------------------------------------- */

/* a * p with a zero adjoint contributing exactly 0 (ft_backward skips such
   nodes), on every row of a pack.  It is the legacy multiply (+0 whenever an
   operand is +-0, Inf and NaN included) plus a * 0 folded into the row's
   finiteness accumulator.  Against a guarded a * p this differs only

   - for a non-finite adjoint: ft_backward's products of it are all
     non-finite, so the row is flagged; here the a * 0 term flags it;
   - in the sign of zero results (a != 0, p = +-0), which reach only Gram
     sums starting from +0, zero adjoints and the accumulator;

   so the assembled system and every row flag are the same. */
__device__ float vm_fmul_legacy(float a, float b) __asm("llvm.amdgcn.fmul.legacy");
/* The adjoint product of one row: f32 rows use the legacy multiply; f64 rows
   (the plain FP64 arm) select +0 for a zero adjoint, which is the same
   contract (an Inf or NaN partial under a zero adjoint contributes 0). */
__device__ __forceinline__ float vm_mul0(float a, float p) { return vm_fmul_legacy(a, p); }
__device__ __forceinline__ double vm_mul0(double a, double p) { return a == 0.0 ? 0.0 : a * p; }
template <typename V> __device__ __forceinline__ V vm_zmul(V a, V p, V &fin)
{
    typedef typename VmScalar<V>::T T;
    V r;
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) {
        r[q] = vm_mul0(a[q], p[q]);
        fin[q] += a[q] * (T)0;
    }
    return r;
}
template <typename V> __device__ __forceinline__ V vm_zmul(V a, typename VmScalar<V>::T c, V &fin)
{
    typedef typename VmScalar<V>::T T;
    V r;
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) {
        r[q] = vm_mul0(a[q], c);
        fin[q] += a[q] * (T)0;
    }
    return r;
}

/* Gradient column write of the pack's rows (rows lane + 64 q of the block,
   in the pack's element type) plus the finiteness accumulator (0 while
   every gradient is finite, NaN after). */
template <typename V> __device__ __forceinline__ void vm_gwrite(char *xl, uint32_t x, V g, V &fin)
{
    typedef typename VmScalar<V>::T T;
    T *p = (T *)(xl + (x >> VMR_GOFS_SHIFT));
#pragma unroll
    for (int q = 0; q < VmScalar<V>::R; ++q) {
        p[64 * q] = g[q];
        fin[q] += g[q] * (T)0;
    }
}

/* Runs the reverse stream seeded with `seed` on a pack of rows (f32, or f64
   for the plain FP64 arm).  The inputs are in U slots 0..v-1 (the stream was
   compiled with them there); gradient columns go to the LDS block through xl
   (this lane's byte address of row 0, column 0; the stream's column offsets
   were compiled for that block's stride and element size).  Returns the
   finiteness accumulator. */
template <typename V, typename FU, typename FB, typename FA>
__device__ __forceinline__ V vm_reverse(const uint2 *__restrict__ rc, int nr, V seed,
                                        const typename VmScalar<V>::T *__restrict__ cst,
                                        const FU &U0, const FU &U1, const FB &B0, const FB &B1,
                                        FA &A0, FA &A1, char *xl)
{
    typedef VmPack<V> P;
    typedef typename VmScalar<V>::T T;
    V a = seed, fin = vm_bcast<V>((T)0);
    const uint32_t end = 8u * (uint32_t)nr;
    for (uint32_t o = 0; o < end; o += 8u) {
        const uint2 w = vm_word(rc, o);
        const uint32_t x = w.x, y = w.y;
        /* One handler bit per word, tested in frequency order. */
        if (x & VMR_PUSH) {
            P::set(A0, A1, A1, A1, VMR_Y_AS(y), a);
        } else if (x & VMR_VMC) {
            vm_gwrite<V>(xl, x, vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_XV(y)), fin), fin);
            a = P::get(A0, A1, A1, A1, VMR_Y_AR(y));
        } else if (x & VMR_UNMC_N) {
            vm_gwrite<V>(xl, x, vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin), fin);
            a = vm_zmul<V>(vm_zmul<V>(a, vm_const(cst, y), fin), P::get(B0, B1, B1, B1, VMR_Y_BT(y)), fin);
        } else if (x & VMR_MULV) {
            a = vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_XV(y)), fin);
        } else if (x & VMR_ADDC_N) {
            vm_gwrite<V>(xl, x, a, fin);
        } else if (x & VMR_MULC_N) {
            vm_gwrite<V>(xl, x, vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin), fin);
            a = vm_zmul<V>(a, vm_const(cst, y), fin);
        } else if (x & VMR_MULC_P) {
            vm_gwrite<V>(xl, x, vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin), fin);
            a = P::get(A0, A1, A1, A1, VMR_Y_AR(y));
        } else if (x & VMR_UNMC_P) {
            vm_gwrite<V>(xl, x, vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin), fin);
            a = P::get(A0, A1, A1, A1, VMR_Y_AR(y));
        } else if (x & VMR_UN) {
            a = vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin);
        } else if (x & VMR_ADDC_P) {
            vm_gwrite<V>(xl, x, a, fin);
            a = P::get(A0, A1, A1, A1, VMR_Y_AR(y));
        } else {
            /* R_GEN: two word pairs. */
            o += 8u;
            const uint32_t x2 = vm_word(rc, o).x;
            const unsigned op = VMR_Y_OP(y);
            V gl, gr;
            if (x & VMR_UNARY) {
                gl = vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin);
                gr = gl;
            } else if (op == OP_ADD) {
                gl = a;
                gr = a;
            } else if (op == OP_SUB) {
                gl = a;
                gr = -a;
            } else {
                gl = vm_zmul<V>(a, P::get(U0, U1, U1, U1, VMR_Y_UT(y)), fin);
                gr = vm_zmul<V>(a, P::get(B0, B1, B1, B1, VMR_Y_BT(y)), fin);
            }
            if (x & VMR_LCONST) vm_gwrite<V>(xl, x, gl, fin);
            if (x & VMR_RCONST) vm_gwrite<V>(xl, x2, gr, fin);
            if ((x & VMR_LNODE) && (x & VMR_RNODE)) {
                P::set(A0, A1, A1, A1, VMR_Y_AS(y), gl);
                a = gr;
            } else if (x & VMR_LNODE) {
                a = gl;
            } else if (x & VMR_RNODE) {
                a = gr;
            } else {
                a = P::get(A0, A1, A1, A1, VMR_Y_AR(y));
            }
        }
    }
    return fin;
}

#endif
