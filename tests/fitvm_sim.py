"""Reference simulator of the fit VM, used by the CPU tests only.

It executes the word pairs produced by rch_fitvm_compile with the semantics
documented in csrc/rc_fitvm.h (the constants are parsed from that header, so
the simulator follows the encoding the device code is built from), on numpy
rows in float64.  Besides the forward value and the constant gradients it
records every register-file read and write, so a test can check the slot
discipline: no read of a slot this pass has not written, no read of a sink.
"""
from __future__ import annotations

import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "the-rule-and-compass_amd", "csrc", "rc_fitvm.h")


def _defines():
    out = {}
    for line in open(HEADER):
        m = re.match(r"#define\s+(VM\w+)\s+\(1u << (\d+)\)", line)
        if m:
            out[m.group(1)] = 1 << int(m.group(2))
            continue
        m = re.match(r"#define\s+(VM\w+)\s+(\d+)u?\s*(/\*.*)?$", line)
        if m:
            out[m.group(1)] = int(m.group(2))
    return out


D = _defines()

OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_POW = 2, 3, 4, 5, 20


def _unary(op, a):
    with np.errstate(all="ignore"):
        table = {
            6: np.sin, 7: np.cos, 8: np.tan, 9: lambda v: np.where(v > 700, np.inf, np.exp(v)),
            10: np.log, 11: np.sqrt, 12: np.abs, 13: np.tanh,
            14: lambda v: 1.0 / (1.0 + np.exp(-v)), 15: lambda v: v * v,
            16: lambda v: v * v * v, 17: lambda v: -v, 18: lambda v: np.exp(-v * v),
            21: np.sinh, 22: np.cosh,
        }
        if op in table:
            return table[op](a)
        import math
        if op == 28:
            return np.vectorize(math.erf)(a)
        if op == 29:
            return np.vectorize(math.erfc)(a)
    raise ValueError(op)


def _partial(op, x, f):
    with np.errstate(all="ignore"):
        if op == 17:
            return -np.ones_like(x)
        if op == 9:
            return f
        if op == 10:
            return 1.0 / x
        if op == 12:
            return np.where(x >= 0, 1.0, -1.0)
        if op == 11:
            return 1.0 / (2.0 * f)
        if op == 6:
            return np.cos(x)
        if op == 7:
            return -np.sin(x)
        if op == 8:
            c = np.cos(x)
            return 1.0 / (c * c)
        if op == 21:
            return np.cosh(x)
        if op == 22:
            return np.sinh(x)
        if op == 13:
            return 1.0 - f * f
        if op == 14:
            return f * (1.0 - f)
        if op == 18:
            return -2.0 * x * f
        if op == 15:
            return 2.0 * x
        if op == 16:
            return 3.0 * x * x
        if op == 28:
            return 1.1283791670955126 * np.exp(-x * x)
        if op == 29:
            return -1.1283791670955126 * np.exp(-x * x)
    raise ValueError(op)


def _binary(op, a, b):
    with np.errstate(all="ignore"):
        if op == OP_ADD:
            return a + b
        if op == OP_SUB:
            return a - b
        if op == OP_MUL:
            return a * b
        if op == OP_DIV:
            return np.where(np.abs(b) < 1e-12, np.nan, a / np.where(b == 0, 1.0, b))
        if op == OP_POW:
            return np.power(a, b)
    raise ValueError(op)


def _bpartials(op, a, b):
    with np.errstate(all="ignore"):
        if op == OP_MUL:
            return b, a
        if op == OP_DIV:
            return 1.0 / b, -a / (b * b)
        if op == OP_POW:
            pos = a > 0
            d0 = np.where(pos, b * np.power(np.where(pos, a, 1.0), b - 1.0), 0.0)
            d1 = np.where(pos, np.power(np.where(pos, a, 1.0), b) * np.log(np.where(pos, a, 1.0)), 0.0)
            return d0, d1
    return np.zeros_like(a), np.zeros_like(a)


class SlotError(AssertionError):
    pass


class _File:
    def __init__(self, name, n, rows, sink):
        self.name, self.sink = name, sink
        self.v = np.zeros((n, rows))
        self.w = np.zeros(n, bool)

    def get(self, i):
        if i == self.sink:
            raise SlotError(f"{self.name}: read of the sink slot {i}")
        if not self.w[i]:
            raise SlotError(f"{self.name}: read of unwritten slot {i}")
        return self.v[i].copy()

    def set(self, i, x):
        self.v[i] = x
        self.w[i] = True


def run(fwd, rev, c, X, col_bytes, flags=0, bslots=None, aslots=None):
    """Forward value and constant gradients of every row of X (rows x vars).
    flags: the compiler options the streams were built with (inputs in U,
    an 8-slot U file).  bslots / aslots: the sizes of the B and A files the
    kernel runs the streams with (the 4-row assembly uses 8 and 2); slot
    numbers are masked to the file size, as the device masks them.

    The simulator writes every tape word's B slot and every general reverse
    word's A slot, NaN where the word keeps nothing (the compiler names the
    sink there), so a word that names a live slot where it should name the
    sink corrupts the result."""
    X = np.asarray(X, np.float64)
    rows = X.shape[0]
    c = np.asarray(c, np.float64)
    small = bool(flags & D["VMC_SMALL_U"])
    usink = D["VM_U_SINK_SMALL"] if small else D["VM_U_SINK"]
    U = _File("U", 8 if small else D["VM_U_SLOTS"], rows, usink)
    in_u = bool(flags & D["VMC_INPUTS_IN_U"])
    if in_u:
        for q in range(X.shape[1]):
            U.set(q, X[:, q])
    nb = bslots or D["VM_B_SLOTS"]
    na = aslots or D["VM_A_SLOTS"]
    B = _File("B", nb, rows, D["VM_B_SINK"] & (nb - 1))
    A = _File("A", na, rows, D["VM_A_SINK"] & (na - 1))
    garbage = np.full(rows, np.nan)
    tos = np.zeros(rows)

    def cval(y):
        off = y & 0xFF
        assert off % 4 == 0
        return c[off // 4]

    def xin(i):
        if in_u:
            return U.get(i)
        return X[:, i]

    fw = np.asarray(fwd, np.uint32).reshape(-1, 2)
    for x, y in fw:
        x, y = int(x), int(y)
        us, ur, bs, xv, op = (y >> 8) & 15, (y >> 12) & 15, (y >> 16) & 15, (y >> 20) & 7, (y >> 23) & 31
        bs &= nb - 1
        if small:
            ur &= 7
        uv = tos
        if x & D["VMF_UNMC"]:
            f = _unary(op, tos)
            B.set(bs, _partial(op, tos, f) if x & D["VMF_NEED"] else garbage)
            uv = f
            nt = f * cval(y)
        elif x & D["VMF_VMC"]:
            nt = xin(xv) * cval(y)
        elif x & D["VMF_ADD"]:
            nt = U.get(ur) + tos
        elif x & D["VMF_MULC"]:
            nt = tos * cval(y)
        elif x & D["VMF_PUSHV"]:
            nt = xin(xv).copy()
        elif x & D["VMF_UN"]:
            f = _unary(op, tos)
            if x & D["VMF_NEED"]:
                uv = _partial(op, tos, f)
            nt = f
        elif x & D["VMF_MULV"]:
            nt = tos * xin(xv)
        elif x & D["VMF_ADDC"]:
            nt = tos + cval(y)
        elif x & D["VMF_PUSHC"]:
            nt = np.full(rows, cval(y))
        elif x & D["VMF_BIN"]:
            if x & D["VMF_RC"]:
                l, r = tos, np.full(rows, cval(y))
            elif x & D["VMF_RV"]:
                l, r = tos, xin(xv)
            else:
                l, r = U.get(ur), tos
            nt = _binary(op, l, r)
            d1 = garbage
            if x & (D["VMF_D0"] | D["VMF_D1"]):
                d0, d1p = _bpartials(op, l, r)
                uv = d0
                if x & D["VMF_D1"]:
                    d1 = d1p
            B.set(bs, d1)
        else:
            raise AssertionError(f"forward word without a handler: {x:#x}")
        if small:
            us &= 7
        if in_u and us < X.shape[1]:
            raise SlotError(f"U: write into input slot {us}")
        U.v[us] = uv
        if us != usink:
            U.w[us] = True
        if x & D["VMF_ACC"]:
            # after the word's own U write: it may have spilled the left operand
            nt = U.get(ur) + nt
        tos = nt
    pred = tos

    grads = {}

    def gw(xword, g):
        off = xword >> D["VMR_GOFS_SHIFT"]
        assert off % col_bytes == 0
        col = off // col_bytes
        assert col not in grads, f"gradient column {col} written twice"
        grads[col] = g

    def z(a, p):
        with np.errstate(all="ignore"):
            return np.where(a == 0, 0.0, a * p)

    a = np.ones(rows)
    rw = np.asarray(rev, np.uint32).reshape(-1, 2)
    i = 0
    while i < len(rw):
        x, y = int(rw[i][0]), int(rw[i][1])
        ut, bt, as_, ar, xv, op = ((y >> 8) & 15, (y >> 12) & 15, (y >> 16) & 7, (y >> 19) & 7,
                                   (y >> 22) & 7, (y >> 25) & 31)
        bt &= nb - 1
        as_ &= na - 1
        ar &= na - 1

        def pop():
            return A.get(ar) if ar != A.sink else np.full(rows, np.nan)

        if x & D["VMR_UNMC_N"]:
            gw(x, z(a, U.get(ut)))
            a = z(z(a, cval(y)), B.get(bt))
        elif x & D["VMR_VMC"]:
            gw(x, z(a, xin(xv)))
            a = pop()
        elif x & D["VMR_PUSH"]:
            A.set(as_, a)
        elif x & D["VMR_MULC_N"]:
            gw(x, z(a, U.get(ut)))
            a = z(a, cval(y))
        elif x & D["VMR_UNMC_P"]:
            gw(x, z(a, U.get(ut)))
            a = pop()
        elif x & D["VMR_UN"]:
            a = z(a, U.get(ut))
        elif x & D["VMR_MULC_P"]:
            gw(x, z(a, U.get(ut)))
            a = pop()
        elif x & D["VMR_MULV"]:
            a = z(a, xin(xv))
        elif x & D["VMR_ADDC_N"]:
            gw(x, a.copy())
        elif x & D["VMR_ADDC_P"]:
            gw(x, a.copy())
            a = pop()
        elif x & D["VMR_GEN"]:
            x2 = int(rw[i + 1][0])
            i += 1
            lneed = bool(x & (D["VMR_LNODE"] | D["VMR_LCONST"]))
            rneed = bool(x & (D["VMR_RNODE"] | D["VMR_RCONST"]))
            gl = gr = None
            if x & D["VMR_UNARY"]:
                gl = z(a, U.get(ut))
            elif op == OP_ADD:
                gl, gr = a.copy(), a.copy()
            elif op == OP_SUB:
                gl, gr = a.copy(), -a
            else:
                # only the partials of the children the sweep needs exist
                gl = z(a, U.get(ut)) if lneed else None
                gr = z(a, B.get(bt)) if rneed else None
            if x & D["VMR_LCONST"]:
                gw(x, gl)
            if x & D["VMR_RCONST"]:
                gw(x2, gr)
            ln, rn = bool(x & D["VMR_LNODE"]), bool(x & D["VMR_RNODE"])
            A.set(as_, gl if (ln and rn) else garbage)
            if ln and rn:
                a = gr
            elif ln:
                a = gl
            elif rn:
                a = gr
            else:
                a = pop()
        else:
            raise AssertionError(f"reverse word without a handler: {x:#x}")
        i += 1
    return pred, grads
