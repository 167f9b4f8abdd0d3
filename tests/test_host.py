"""CPU tests of the boundary library and the host-side program builders.

No compute call needs a GPU here: the library must load, export every symbol
the public headers declare, and its pure-host builders (compile_program,
gpu_fit_plan) must emit exactly the words the reference host emits
(gpu.rs:198-438, fit.rs:985-1196).
"""
import os
import re
import struct

import numpy as np
import pytest

from rcsr_hip import host, lib
from rcsr_hip.expr import (ACOSH, ADD, ASINH, ATANH, COS, DIV, EXP, HALT, IF, LOG, MUL, NEG,
                           SIN, SQRT, STEP, SUB, ABS, MIN, const, node, var)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    names = set()
    for h in ("rcsr_hip.h", "rcsr_host.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(rch?_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    L = lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 30
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared <= set(lib.SIGNATURES), sorted(declared - set(lib.SIGNATURES))


def test_fit_kernel_info_refuses_more_than_twenty_constants():
    import ctypes as C
    L = lib.load()
    t, s, m = C.c_int(), C.c_int(), C.c_int()
    assert L.rc_fit_kernel_info(21, C.byref(t), C.byref(s), C.byref(m)) == -1


def f32(v):
    return struct.unpack("<I", struct.pack("<f", v))[0]


def test_compile_program_emits_reference_postfix():
    # 2*x0 + sin(x1)
    t = node(ADD, node(MUL, const(2.0), var(0)), node(SIN, var(1)))
    w = host.compile_program(t).tolist()
    assert w == [0, f32(2.0), 1, 0, 4, 1, 1, 6, 2, HALT]


def test_compile_program_lowers_inverse_hyperbolics():
    for op in (ASINH, ACOSH, ATANH):
        w = host.compile_program(node(op, var(0))).tolist()
        assert w[-1] == HALT
        # The lowered program uses only native opcodes (gpu.rs:320-345):
        # walk the words, skipping each CONST/VAR operand word.
        i = 0
        while i < len(w):
            assert w[i] <= 42 or w[i] == HALT, w[i]
            i += 2 if w[i] in (0, 1) else 1
    # asinh lowers to sgn * ln(|x| + sqrt(x*x + 1)) with sgn = step(x) - step(-x)
    w = host.compile_program(node(ASINH, var(0))).tolist()
    assert STEP in w and LOG in w and SQRT in w and ABS in w


def test_compile_program_rejects_deep_stacks_and_unknown_ops():
    deep = var(0)
    for _ in range(60):
        deep = node(ADD, var(0), deep)   # right-leaning: depth grows by one per level
    with pytest.raises(host.CompileError) as e:
        host.compile_program(deep)
    assert e.value.kind == "stack_overflow" and e.value.peak > 56
    shallow = var(0)
    for _ in range(60):
        shallow = node(ADD, shallow, var(0))  # left-leaning: depth 2
    assert host.compile_program(shallow)[-1] == HALT


def test_fit_plan_records_are_post_order_with_slots():
    # c0 * x0 + c1  ->  [CONST s0][VAR 0][MUL 0 1][CONST s1][ADD 2 3]
    t = node(ADD, node(MUL, const(0.5), var(0)), const(-1.25))
    p = host.gpu_fit_plan(t)
    assert p.nconst == 2 and p.nodes == 5
    assert p.c0.tolist() == [0.5, -1.25]
    rec = p.words.reshape(-1, 4).tolist()
    U = 0xFFFFFFFF
    assert rec == [[0, 0, 0, 0], [1, 0, 0, 0], [4, 0, 1, U], [0, 0, 0, 1], [2, 2, 3, U]]


def test_fit_plan_rejections_match_the_reference_labels():
    assert host.gpu_fit_plan_reason(node(SIN, var(0))) == host.GpuFitReject.NO_CONSTANTS
    many = node(MUL, const(0.1), var(0))
    for k in range(1, 21):
        many = node(ADD, many, node(MUL, const(0.2 + 0.01 * k), var(0)))
    assert host.gpu_fit_plan_reason(many) == host.GpuFitReject.TOO_MANY_CONSTANTS
    twenty = node(MUL, const(0.1), var(0))
    for k in range(1, 20):
        twenty = node(ADD, twenty, node(MUL, const(0.2), var(0)))
    assert host.gpu_fit_plan(twenty).nconst == 20
    assert host.gpu_fit_plan_reason(node(MIN, const(1.0), var(0))) == \
        host.GpuFitReject.UNSUPPORTED_OP
    big = const(1.0)
    for _ in range(50):
        big = node(ADD, big, node(NEG, var(0)))
    assert host.gpu_fit_plan_reason(big) == host.GpuFitReject.TOO_MANY_NODES


def test_status_usability_matches_the_reference():
    assert [s for s in range(10) if host.fit_status_usable(s)] == [0, 4, 5, 6, 7, 8]


def test_workloads_are_seeded_and_shaped():
    from rcsr_hip import workloads
    a = workloads.config1()
    b = workloads.config1()
    assert a.x.shape == (1000, 1) and len(a.trees) == 800
    assert np.array_equal(a.x, b.x) and str(a.trees[17]) == str(b.trees[17])
    assert all(len(t.constants()) <= 4 for t in a.trees)
    c3 = workloads.config3(n=4096, ncand=64)
    ks = [len(t.constants()) for t in c3.trees]
    assert min(ks) >= 1 and max(ks) <= 20
    assert all(host.gpu_fit_plan(t) is not None for t in c3.trees)
    assert np.all(np.isfinite(c3.y))


def _graph_key(code_words, nrowsets, nprog=64, ncands=64, nitems=128, nlist=256, kmax=8, tri=45):
    import ctypes as C
    L = lib.load()
    buf = (C.c_ulonglong * 64)()
    n = L.rch_fit_graph_key(code_words, ncands, nrowsets, nitems, nlist, nprog, kmax, tri, buf, 64)
    assert n > 0
    return tuple(buf[:n])


def test_fit_graph_key_separates_regions_that_end_at_the_same_byte():
    # Layouts whose region ends at the same byte (a shorter code block and one
    # more row set) must not share a replay key: the graph bakes in the
    # offsets of the candidate and row-set tables.
    fields = 22                      # region offsets after the 7 shape words
    dev_end_at = 7 + fields - 1
    seen = {}
    collisions = 0
    for code_words in range(64, 64 * 40, 16):
        for nrowsets in range(1, 18):
            k = _graph_key(code_words, nrowsets)
            offsets = k[7:7 + fields]
            shape = (k[:7], k[dev_end_at])
            for other in seen.get(shape, []):
                if other[7:7 + fields] != offsets:
                    collisions += 1
                    assert other != k
            seen.setdefault(shape, []).append(k)
    # The shape words and region end alone would have replayed a stale graph.
    assert collisions > 0
    # Equal layouts give equal keys.
    assert _graph_key(1000, 3) == _graph_key(1000, 3)
