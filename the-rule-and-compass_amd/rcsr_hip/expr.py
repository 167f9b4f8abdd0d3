"""Expression trees for the host side of the device arm.

Mirrors the reference engine's expression core (my-version/src/expr.rs:12-77
for `Op`, 526-541 for `Node`) only as far as the device arm needs it: a tree
of operators, constants and input columns that the bytecode compiler
(`compile_program`, gpu.rs:198) and the fit-plan builder
(`gpu_fit_plan_reason`, fit.rs:985) consume.

Across every C ABI of this project a tree travels as a pre-order node list:
`ops[i]` is the operator code and `vals[i]` the constant value (for CONST) or
the column index (for VAR).  Operator codes are the device opcodes of
eval.cu:34-83 (= gpu.rs:29-73) for every operator the device knows, plus
ASINH/ACOSH/ATANH, which the bytecode compiler lowers (gpu.rs:320-345).
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

# Device opcodes (eval.cu:34-83, gpu.rs:29-73).
CONST, VAR, ADD, SUB, MUL, DIV = 0, 1, 2, 3, 4, 5
SIN, COS, TAN, EXP, LOG, SQRT, ABS, TANH, SIG = 6, 7, 8, 9, 10, 11, 12, 13, 14
SQUARE, CUBE, NEG, GAUSS, STEP, POW, SINH, COSH = 15, 16, 17, 18, 19, 20, 21, 22
FLOOR, CEIL, ROUND, MOD, ATAN2, ERF, ERFC, MIN, MAX, IF = 23, 24, 25, 26, 27, 28, 29, 30, 31, 32
SIGN, EQUAL, LESS, LESS_OR_EQUAL, GREATER, GREATER_OR_EQUAL = 33, 34, 35, 36, 37, 38
AND, OR, XOR, NOT = 39, 40, 41, 42
HALT = 255
# Tree-only operators the compiler lowers before emission.
ASINH, ACOSH, ATANH = 43, 44, 45

BINARY = {ADD, SUB, MUL, DIV, POW, MOD, ATAN2, MIN, MAX, EQUAL, LESS, LESS_OR_EQUAL,
          GREATER, GREATER_OR_EQUAL, AND, OR, XOR}
TERNARY = {IF}
LEAF = {CONST, VAR}

NAMES = {
    CONST: "const", VAR: "var", ADD: "+", SUB: "-", MUL: "*", DIV: "/", SIN: "sin",
    COS: "cos", TAN: "tan", EXP: "exp", LOG: "log", SQRT: "sqrt", ABS: "abs",
    TANH: "tanh", SIG: "logistic", SQUARE: "square", CUBE: "cube", NEG: "neg",
    GAUSS: "gauss", STEP: "step", POW: "pow", SINH: "sinh", COSH: "cosh",
    FLOOR: "floor", CEIL: "ceil", ROUND: "round", MOD: "mod", ATAN2: "atan2",
    ERF: "erf", ERFC: "erfc", MIN: "min", MAX: "max", IF: "if", SIGN: "sign",
    EQUAL: "==", LESS: "<", LESS_OR_EQUAL: "<=", GREATER: ">",
    GREATER_OR_EQUAL: ">=", AND: "and", OR: "or", XOR: "xor", NOT: "not",
    ASINH: "asinh", ACOSH: "acosh", ATANH: "atanh",
}


def arity(op: int) -> int:
    if op in LEAF:
        return 0
    if op in BINARY:
        return 2
    if op in TERNARY:
        return 3
    if 0 <= op <= ATANH:
        return 1
    raise ValueError(f"unknown operator code {op}")


@dataclass
class Node:
    op: int
    val: float = 0.0
    kids: List["Node"] = field(default_factory=list)

    def preorder(self) -> Tuple[np.ndarray, np.ndarray]:
        ops: List[int] = []
        vals: List[float] = []

        def walk(n: "Node") -> None:
            ops.append(n.op)
            vals.append(float(n.val))
            for k in n.kids:
                walk(k)

        walk(self)
        return np.asarray(ops, dtype=np.int32), np.asarray(vals, dtype=np.float64)

    def constants(self) -> List[float]:
        """Constant values in slot order (left-to-right leaves)."""
        out: List[float] = []

        def walk(n: "Node") -> None:
            if n.op == CONST:
                out.append(n.val)
            for k in n.kids:
                walk(k)

        walk(self)
        return out

    def with_constants(self, values: Sequence[float]) -> "Node":
        it = iter(values)

        def walk(n: "Node") -> "Node":
            if n.op == CONST:
                return Node(CONST, float(next(it)))
            return Node(n.op, n.val, [walk(k) for k in n.kids])

        return walk(self)

    def size(self) -> int:
        return 1 + sum(k.size() for k in self.kids)

    def __str__(self) -> str:
        if self.op == CONST:
            return repr(self.val)
        if self.op == VAR:
            return f"x{int(self.val)}"
        args = ", ".join(str(k) for k in self.kids)
        return f"{NAMES.get(self.op, self.op)}({args})"


def const(v: float) -> Node:
    return Node(CONST, float(v))


def var(i: int) -> Node:
    return Node(VAR, float(i))


def node(op: int, *kids: Node) -> Node:
    if len(kids) != arity(op):
        raise ValueError(f"{NAMES.get(op, op)} takes {arity(op)} operands, got {len(kids)}")
    return Node(op, 0.0, list(kids))


def from_preorder(ops: Sequence[int], vals: Sequence[float]) -> Node:
    pos = 0

    def walk() -> Node:
        nonlocal pos
        op = int(ops[pos])
        v = float(vals[pos])
        pos += 1
        return Node(op, v, [walk() for _ in range(arity(op))])

    t = walk()
    if pos != len(ops):
        raise ValueError("trailing nodes in pre-order list")
    return t


def pack_trees(trees: Sequence[Node]) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Concatenate pre-order lists; returns (ops, vals, offsets[ntree+1])."""
    ops, vals, off = [], [], [0]
    for t in trees:
        o, v = t.preorder()
        ops.append(o)
        vals.append(v)
        off.append(off[-1] + len(o))
    return (np.concatenate(ops).astype(np.int32) if ops else np.zeros(0, np.int32),
            np.concatenate(vals) if vals else np.zeros(0),
            np.asarray(off, dtype=np.int32))


def f32_bits(v: float) -> int:
    return struct.unpack("<I", struct.pack("<f", v))[0]
