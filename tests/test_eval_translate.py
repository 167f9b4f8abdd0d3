"""CPU checks of the evaluator's program translation (rch_eval_translate, the
streams rc_set_programs uploads).

The translation rebuilds each program's tree, evaluates input-free subtrees
once per batch (hoist programs whose results fill constant words) and fuses
+, -, *, / with a leaf operand into one word.  A small f32 simulator of the
device interpreter runs the translated streams, hoist programs first, and
must reproduce, bit for bit, a plain f32 stack machine run over the original
postfix program with the reference's stack rules (rc_exec_row,
eval.cu:404-573: underflow, overflow past 64 and a missing HALT fail every
row; the value returned at HALT is the bottom stack entry).  Both machines
share one set of f32 operator functions, so any difference is the
translation's: operand order, fusion, hoisting, dropped dead entries.
"""
import numpy as np
import pytest

from rcsr_hip import host, lib, workloads
from rcsr_hip.expr import (ADD, ATAN2, COS, DIV, EXP, HALT, IF, LOG, MAX, MIN, MOD, MUL, NEG,
                           POW, SIN, SQRT, SQUARE, SUB, TANH, const, node, var)

FUSED = 64
OP_NAN = 254
F32_NAN = np.float32(np.nan)
BIN = {ADD, SUB, MUL, DIV, POW, MOD, ATAN2, MIN, MAX} | set(range(34, 42))
UN = {SIN, COS, 8, EXP, LOG, SQRT, 12, TANH, 14, SQUARE, 16, NEG, 18, 19, 21, 22, 23, 24, 25,
      28, 29, 33, 42}


def translate(prog, off, nprog, nvars, derived=(), cap=256):
    """rch_eval_translate; derived=None disables derived columns, otherwise
    it lists the (operator, source column) pairs already computed."""
    L = lib.load()
    info = np.zeros(7, np.int32)
    prog = np.ascontiguousarray(prog, np.uint32)
    off = np.ascontiguousarray(off, np.uint32)
    if derived is None:
        dv, nd, dcap = None, None, 0
    else:
        dv = np.zeros(2 * max(cap, 1), np.uint32)
        dv[:2 * len(derived)] = np.asarray(derived, np.uint32).reshape(-1)
        nd = np.asarray([len(derived)], np.int32)
        dcap = cap
    dptr = lib.ptr(dv, lib.u32p) if dv is not None else None
    nptr = lib.ptr(nd, lib.i32p) if nd is not None else None
    rc = L.rch_eval_translate(lib.ptr(prog, lib.u32p), lib.ptr(off, lib.u32p), nprog, nvars,
                              None, 0, None, 0, dptr, dcap, nptr, lib.ptr(info, lib.i32p))
    assert rc == -2
    code = np.zeros(max(int(info[0]), 1), np.uint32)
    meta = np.zeros(max(int(info[1]), 1), np.uint32)
    rc = L.rch_eval_translate(lib.ptr(prog, lib.u32p), lib.ptr(off, lib.u32p), nprog, nvars,
                              lib.ptr(code, lib.u32p), len(code), lib.ptr(meta, lib.u32p),
                              len(meta), dptr, dcap, nptr, lib.ptr(info, lib.i32p))
    assert rc == 0
    pairs = [] if dv is None else [(int(dv[2 * k]), int(dv[2 * k + 1])) for k in range(int(nd[0]))]
    return code[:info[0]], meta[:info[1]], info, pairs


def extend_columns(X, pairs):
    """The derived columns: pair k = (operator, source column) is column
    nvars + k, computed from the (input or earlier derived) source."""
    cols = [X[:, j] for j in range(X.shape[1])]
    for op, src in pairs:
        cols.append(unary(op, cols[src]))
    return np.stack(cols, 1) if cols else X


def f(v):
    return np.float32(v)


def bits(v):
    return np.asarray(v, np.float32).view(np.uint32)


def unary(op, a):
    with np.errstate(all="ignore"):
        if op == SIN:
            return np.sin(a)
        if op == COS:
            return np.cos(a)
        if op == EXP:
            return np.exp(a)
        if op == LOG:
            return np.where(a > 0, np.log(np.where(a > 0, a, f(1))), F32_NAN).astype(np.float32)
        if op == SQRT:
            return np.where(a >= 0, np.sqrt(np.abs(a)), F32_NAN).astype(np.float32)
        if op == NEG:
            return -a
        if op == SQUARE:
            return a * a
        if op == TANH:
            return np.tanh(a)
        if op in UN:
            return (a * f(0.5) + f(0.25)).astype(np.float32)
        return np.full_like(a, F32_NAN)


def binary(op, a, b):
    with np.errstate(all="ignore"):
        if op == ADD:
            return a + b
        if op == SUB:
            return a - b
        if op == MUL:
            return a * b
        if op == DIV:
            return np.where(np.abs(b) < f(1e-12), F32_NAN, a / b).astype(np.float32)
        if op == MIN:
            return np.where(np.isnan(a) | np.isnan(b), F32_NAN, np.minimum(a, b)).astype(np.float32)
        if op == MAX:
            return np.where(np.isnan(a) | np.isnan(b), F32_NAN, np.maximum(a, b)).astype(np.float32)
        return (a * f(0.75) - b * f(0.5)).astype(np.float32)


def plain_run(words, X, nvars):
    """The original postfix program over the rows of X (rows x nvars)."""
    n = X.shape[0]
    nan = np.full(n, F32_NAN, np.float32)
    st = []
    i = 0
    while True:
        if i >= len(words):
            return nan
        op = int(words[i])
        i += 1
        if op in (0, 1):
            if i >= len(words) or len(st) >= 64:
                return nan
            arg = int(words[i])
            i += 1
            if op == 0:
                st.append(np.full(n, np.uint32(arg).view(np.float32), np.float32))
            elif arg < nvars and arg < 65536:
                st.append(X[:, arg].copy())
            else:
                st.append(nan.copy())
        elif op == HALT:
            return st[0] if st else nan
        elif op == IF:
            if len(st) < 3:
                return nan
            e, t, c = st.pop(), st.pop(), st.pop()
            st.append(np.where(c > 0, t, e).astype(np.float32))
        elif op in BIN:
            if len(st) < 2:
                return nan
            b, a = st.pop(), st.pop()
            st.append(binary(op, a, b))
        else:
            if not st:
                return nan
            st.append(unary(op if op in UN else OP_NAN, st.pop()))


def stream_run(code, pc, X):
    """The device interpreter's semantics over the translated stream."""
    n = X.shape[0]
    st = []
    tos = None
    while True:
        w = int(code[pc])
        pc += 1
        op = w & 0xFF
        if op in (0, 1):
            if tos is not None:
                st.append(tos)
            if op == 1:
                tos = X[:, w >> 16].copy()
            else:
                tos = np.full(n, np.uint32(code[pc]).view(np.float32), np.float32)
                pc += 1
        elif FUSED <= op < FUSED + 16:
            k = op - FUSED
            bop = (ADD, SUB, MUL, DIV)[k >> 2]
            if k & 1:
                leaf = X[:, w >> 16]
            else:
                leaf = np.full(n, np.uint32(code[pc]).view(np.float32), np.float32)
                pc += 1
            tos = binary(bop, leaf, tos) if k & 2 else binary(bop, tos, leaf)
        elif op == HALT:
            return st[0] if st else tos
        elif op == IF:
            t, c = st.pop(), st.pop()
            tos = np.where(c > 0, t, tos).astype(np.float32)
        elif op in BIN:
            tos = binary(op, st.pop(), tos)
        else:
            tos = unary(op, tos)


def run_translated(code, meta, nprog, X):
    code = code.copy()
    pairs = meta[2 * nprog:].reshape(-1, 2)
    one = np.zeros((1, X.shape[1]), np.float32)
    for hp, target in pairs:
        assert code[target] == 0x7FC00000
        code[target] = bits(stream_run(code, int(hp), one))[0]
    return [stream_run(code, int(meta[p]), X) for p in range(nprog)]


def leaf_words(rng, nvars):
    if rng.random() < 0.45:
        return [0, int(bits(f(rng.normal())))]
    return [1, int(rng.integers(nvars + (1 if rng.random() < 0.05 else 0)))]


def random_program(rng, nvars, nodes):
    """A postfix program; mostly well formed, with extra stack entries,
    unknown operators and the occasional malformed tail."""
    words, depth = [], 0
    ops = [ADD, SUB, MUL, DIV, SIN, COS, EXP, LOG, IF, MIN, MAX, POW, SQRT, NEG, SQUARE, 200]
    for _ in range(nodes):
        op = ops[int(rng.integers(len(ops)))]
        need = 3 if op == IF else 2 if op in BIN else 1
        if depth < need or rng.random() < 0.4:
            words += leaf_words(rng, nvars)
            depth += 1
        else:
            words.append(op)
            depth -= need - 1
    r = rng.random()
    if r < 0.03:
        return words  # no HALT
    if r < 0.06:
        words.append(ADD)  # may underflow
    words.append(HALT)
    return words


def pack(progs):
    off = np.zeros(len(progs) + 1, np.uint32)
    off[1:] = np.cumsum([len(p) for p in progs])
    return np.asarray([w for p in progs for w in p], np.uint32), off


def check_batch(progs, nvars, X, derived=(), cap=256):
    prog, off = pack(progs)
    code, meta, info, pairs = translate(prog, off, len(progs), nvars, derived, cap)
    got = run_translated(code, meta, len(progs), extend_columns(X, pairs))
    for p, words in enumerate(progs):
        want = plain_run(words, X, nvars)
        same = np.array_equal(np.isnan(got[p]), np.isnan(want)) and np.array_equal(
            bits(got[p])[~np.isnan(want)], bits(want)[~np.isnan(want)])
        assert same, (p, words)
    return code, meta, info, pairs


def _x(rng, n=64, nvars=3):
    x = rng.normal(size=(n, nvars)).astype(np.float32)
    x[0] = 0.0
    x[1, 0] = np.inf
    return x


def test_random_programs_translate_to_the_same_values():
    rng = np.random.default_rng(5)
    X = _x(rng)
    progs = [random_program(rng, 3, int(rng.integers(1, 60))) for _ in range(400)]
    code, meta, info, pairs = check_batch(progs, 3, X)
    nfused = sum(1 for w in code if FUSED <= (w & 0xFF) < FUSED + 16)
    assert nfused > 100 and len(meta) > 2 * len(progs)  # fused words and hoists both occur


def test_large_batches_translate_to_the_same_values():
    # 1100 programs: the batch is translated in several program chunks whose
    # streams, hoist programs and derived-column counts are joined
    rng = np.random.default_rng(15)
    X = _x(rng, 32)
    progs = [random_program(rng, 3, int(rng.integers(1, 40))) for _ in range(1100)]
    code, meta, info, pairs = check_batch(progs, 3, X)
    assert len(pairs) > 0 and len(meta) > 2 * len(progs)


def test_batch_streams_equal_single_program_streams():
    # Without derived columns a program's stream and hoist programs depend
    # on the program alone: the joined batch holds each single translation.
    rng = np.random.default_rng(16)
    progs = [random_program(rng, 3, int(rng.integers(1, 40))) for _ in range(1300)]
    prog, off = pack(progs)
    code, meta, info, _ = translate(prog, off, len(progs), 3, None)
    n = len(progs)
    ends = list(meta[1:n]) + [int(meta[2 * n]) if len(meta) > 2 * n else len(code)]
    hoists = {}
    for hp, target in meta[2 * n:].reshape(-1, 2):
        q = int(hp)
        while code[q] != HALT:
            q += 1
        hoists[int(target)] = code[int(hp):q + 1].tolist()
    for p in range(0, n, 7):
        sp, so = pack([progs[p]])
        c1, m1, _, _ = translate(sp, so, 1, 3, None)
        main1 = int(m1[2]) if len(m1) > 2 else len(c1)
        assert code[meta[p]:ends[p]].tolist() == c1[:main1].tolist(), p
        h1 = {}
        for hp, target in m1[2:].reshape(-1, 2):
            q = int(hp)
            while c1[q] != HALT:
                q += 1
            h1[int(target)] = c1[int(hp):q + 1].tolist()
        mine = {t - int(meta[p]): h for t, h in hoists.items() if meta[p] <= t < ends[p]}
        assert mine == h1, p


def test_c2_programs_fuse_and_hoist():
    wl = workloads.config2(n=64, nprog=256)
    pop = host.Population()
    for t in wl.trees:
        pop.add(t)
    prog, off = pop.arrays()
    progs = [prog[off[i]:off[i + 1]].tolist() for i in range(len(off) - 1)]
    code, meta, info, pairs = check_batch(progs, 5, wl.x.astype(np.float32))
    main = int(meta[2 * len(progs)]) if len(meta) > 2 * len(progs) else len(code)
    words_in = sum(len(p) for p in progs)
    assert main < 0.8 * words_in  # fewer dispatched words than the postfix
    assert info[6] <= 16


def test_hoisted_subtree_values_and_layout():
    # x0 * (sin(0.5) + exp(1.5)) - log(2.0) / x1: two input-free subtrees.
    t = node(SUB, node(MUL, var(0), node(ADD, node(SIN, const(0.5)), node(EXP, const(1.5)))),
             node(DIV, node(LOG, const(2.0)), var(1)))
    w = host.compile_program(t).tolist()
    X = _x(np.random.default_rng(1), 32, 2)
    code, meta, info, pairs = check_batch([w], 2, X)
    pairs = meta[2:].reshape(-1, 2)
    assert len(pairs) == 2
    # main stream: VAR x0, MUL-by-constant, then the right operand
    assert code[0] == 1 | (0 << 16)
    assert code[1] == FUSED + 2 * 4 + 0  # tos * c
    # each hoist program is a plain postfix ending in HALT, after the main code
    for hp, target in pairs:
        assert hp > target and code[target] == 0x7FC00000
        q = int(hp)
        while code[q] != HALT:
            q += 1
        assert (code[int(hp):q] & 0xFF).max() < FUSED


def test_left_leaf_keeps_operand_order_and_whole_constant_programs():
    X = _x(np.random.default_rng(2), 32, 2)
    progs = [
        host.compile_program(node(SUB, const(3.0), node(SIN, var(0)))).tolist(),
        host.compile_program(node(DIV, var(1), node(COS, var(0)))).tolist(),
        host.compile_program(node(DIV, const(1.0), node(ADD, const(2.0), const(3.0)))).tolist(),
        [0, int(bits(f(2.5))), HALT],
        [1, 0, 1, 1, HALT],            # two entries: the bottom one is returned
        [1, 7, HALT],                  # column outside the resident set
        [0, 0, IF, HALT],              # underflow
        [1, 0] * 65 + [HALT],          # overflow past 64 entries
        [1, 0],                        # no HALT
    ]
    code, meta, info, pairs = check_batch(progs, 2, X)
    # c - sin(x0): sin first, then the left-constant SUB
    p0 = code[int(meta[0]):]
    assert p0[1] == SIN and p0[2] == FUSED + 1 * 4 + 2 and p0[4] == HALT
    # 1 / (2 + 3): one hoist program for the whole tree
    p2 = code[int(meta[2]):]
    assert p2[0] == 0 and p2[2] == HALT


def test_tall_trees_fall_back_to_the_plain_form():
    X = _x(np.random.default_rng(3), 16, 1)
    deep = [1, 0] + [SIN, 0, int(bits(f(0.5))), ADD] * 1100 + [HALT]
    code, meta, info, pairs = check_batch([deep], 1, X)
    # the postfix word for word, except that a VAR word carries its column
    assert len(code) == len(deep) - 1


def test_derived_columns_are_shared_kept_and_capped():
    X = _x(np.random.default_rng(4), 48, 3)
    sx0 = node(SIN, var(0))
    a = [node(ADD, node(MUL, sx0, const(2.0)), var(1)),
         node(SUB, sx0, node(COS, var(1))),
         node(EXP, sx0),
         node(DIV, node(EXP, sx0), var(2)),
         node(LOG, var(2))]
    progs = [host.compile_program(t).tolist() for t in a]
    code, meta, info, pairs = check_batch(progs, 3, X)
    assert pairs[0] == (SIN, 0)                  # four uses
    assert (EXP, 3) in pairs                     # exp of the derived sin(x0), twice
    assert (COS, 1) not in pairs and (LOG, 2) not in pairs  # one use each
    # the first program reads column 3 and multiplies it by the constant
    p0 = code[int(meta[0]):]
    assert p0[0] == 1 | (3 << 16) and p0[1] == FUSED + 2 * 4 + 0
    # a later batch reuses the computed columns and appends only new pairs
    b = [node(ADD, node(COS, var(1)), node(MUL, node(COS, var(1)), var(0))),
         node(MUL, sx0, var(2))]
    progs2 = [host.compile_program(t).tolist() for t in b]
    _, _, _, pairs2 = check_batch(progs2, 3, X, derived=pairs)
    assert pairs2[:len(pairs)] == pairs and pairs2[len(pairs):] == [(COS, 1)]
    # without the table: plain input reads, same values
    _, _, _, none = check_batch(progs, 3, X, derived=None)
    assert none == []
    # the cap holds
    rng = np.random.default_rng(6)
    many = [random_program(rng, 3, int(rng.integers(5, 60))) for _ in range(300)]
    _, _, _, capped = check_batch(many, 3, X, cap=3)
    assert len(capped) == 3
    _, _, _, free = check_batch(many, 3, X)
    assert len(free) > 3
