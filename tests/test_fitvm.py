"""The fit VM compiler against the oracle, on the CPU.

rch_fitvm_compile turns a candidate's post-order records into the register-
slot streams the mixed-precision assembly runs.  Each compiled program is
executed by the reference simulator of the VM (tests/fitvm_sim.py, float64)
and its prediction and constant gradients are compared row by row with the
oracle's (or_eval / or_gradients: the reference's value_of and adjoint,
fit.rs:1211-1365 and eval.cu:1236-1317).  A wrong slot assignment, a
dropped partial, a swapped operand of SUB / DIV / POW or a missed pop all
change a value or a gradient; the simulator also rejects a read of a slot
the pass has not written and a second write of one gradient column.
"""
import ctypes as C

import numpy as np
import pytest

import fitvm_sim
import oracle
from rcsr_hip import host, lib, workloads
from rcsr_hip.expr import (ABS, ADD, COS, CUBE, DIV, ERF, ERFC, EXP, GAUSS, LOG, MUL, NEG, POW,
                           SIG, SIN, SINH, COSH, SQRT, SQUARE, SUB, TAN, TANH, const, node, var)

COL_BYTES = 130 * 4
UNARY = [SIN, COS, TAN, EXP, LOG, SQRT, ABS, TANH, SIG, SQUARE, CUBE, NEG, GAUSS, SINH, COSH,
         ERF, ERFC]
BINARY = [ADD, SUB, MUL, DIV, POW]


INPUTS_IN_U, VALUE_ONLY, SMALL_U = 1, 2, 4


def compile_tree(tree, nvars=3, flags=0):
    plan = host.gpu_fit_plan(tree)
    assert plan is not None
    L = lib.load()
    rec = np.ascontiguousarray(plan.words, np.uint32)
    fw = np.zeros(2 * 4 * 96, np.uint32)
    rv = np.zeros(2 * 8 * 96, np.uint32)
    info = np.zeros(6, np.int32)
    rc = L.rch_fitvm_compile(lib.ptr(rec, lib.u32p), len(rec), plan.nconst, nvars, COL_BYTES,
                             flags, lib.ptr(fw, lib.u32p), len(fw), lib.ptr(rv, lib.u32p), len(rv),
                             lib.ptr(info, lib.i32p))
    return rc, plan, fw[:2 * info[0]], rv[:2 * info[1]], info


def check_tree(tree, X, rel=1e-9, flags=0):
    rc, plan, fw, rv, info = compile_tree(tree, X.shape[1], flags)
    if rc != 0:
        return rc
    pred, grads = fitvm_sim.run(fw, rv, plan.c0, X, COL_BYTES, flags)
    if flags & VALUE_ONLY:
        assert len(rv) == 0 and not grads
        want = oracle.predict(tree, X)
        fin = np.isfinite(want) & np.isfinite(pred)
        assert np.all(fin | (~np.isfinite(pred) & ~np.isfinite(want))), str(tree)
        assert np.allclose(pred[fin], want[fin], rtol=rel, atol=rel), str(tree)
        assert info[2] <= (7 if flags & SMALL_U else 15)
        return 0
    want = oracle.predict(tree, X)
    # The evaluator semantics of oracle.predict turn some domain edges into
    # NaN where the fit's value_of gives +-inf: both non-finite agree.
    both_bad = ~np.isfinite(pred) & ~np.isfinite(want)
    fin = np.isfinite(want) & np.isfinite(pred)
    assert np.all(both_bad | fin), (str(tree), pred, want)
    assert np.allclose(pred[fin], want[fin], rtol=rel, atol=rel), str(tree)
    assert sorted(grads) == list(range(plan.nconst)), (str(tree), sorted(grads))
    for r in range(X.shape[0]):
        if not np.isfinite(want[r]):
            continue
        g, _ = oracle.gradients(tree, X[r])
        for j in range(plan.nconst):
            gv = grads[j][r]
            if not np.isfinite(g[j]):
                continue
            assert abs(gv - g[j]) <= rel * (1.0 + abs(g[j])) * 10, (str(tree), r, j, gv, g[j])
    if info[3] <= 7 and info[4] <= 1:
        # On an 8-slot B file and a 2-slot A file (the 4-row assembly's) the
        # sinks are slots 7 and 1: the streams still compute the same.
        p4, g4 = fitvm_sim.run(fw, rv, plan.c0, X, COL_BYTES, flags, bslots=8, aslots=2)
        assert np.array_equal(p4, pred, equal_nan=True), str(tree)
        assert sorted(g4) == sorted(grads)
        for j in grads:
            assert np.array_equal(g4[j], grads[j], equal_nan=True), (str(tree), j)
    return 0


def full_tree(rng, depth, nvars=3):
    if depth <= 0 or (depth < 4 and rng.random() < 0.3):
        if rng.random() < 0.45:
            return const(float(rng.uniform(0.3, 1.7)))
        return var(int(rng.integers(nvars)))
    if rng.random() < 0.35:
        return node(UNARY[int(rng.integers(len(UNARY)))], full_tree(rng, depth - 1, nvars))
    return node(BINARY[int(rng.integers(len(BINARY)))], full_tree(rng, depth - 1, nvars),
                full_tree(rng, depth - 1, nvars))


@pytest.fixture(scope="module")
def X():
    rng = np.random.default_rng(7)
    return rng.uniform(0.2, 1.8, size=(24, 3))


EDGE = [
    ("lone constant", lambda: const(1.3)),
    ("c*x", lambda: node(MUL, const(0.7), var(0))),
    ("x*c", lambda: node(MUL, var(1), const(0.7))),
    ("c1*c2", lambda: node(MUL, const(0.7), const(1.1))),
    ("c1+c2", lambda: node(ADD, const(0.7), const(1.1))),
    ("c1-c2", lambda: node(SUB, const(0.7), const(1.1))),
    ("c1/c2", lambda: node(DIV, const(0.7), const(1.1))),
    ("exp(c)", lambda: node(EXP, const(0.4))),
    ("exp(c)*c", lambda: node(MUL, node(EXP, const(0.4)), const(0.9))),
    ("c*sin(c x)", lambda: node(MUL, const(1.2), node(SIN, node(MUL, const(0.7), var(0))))),
    ("c*cos(c x + c)", lambda: node(MUL, const(1.2), node(COS, node(ADD, node(MUL, const(0.7), var(0)),
                                                                     const(0.3))))),
    ("c*x/(c + c*x*x)", lambda: node(MUL, const(0.8), node(DIV, var(0), node(
        ADD, const(1.5), node(MUL, node(MUL, const(0.1), var(0)), var(0)))))),
    ("c/(x+c)", lambda: node(DIV, const(2.0), node(ADD, var(1), const(0.5)))),
    ("x-c", lambda: node(SUB, var(0), const(0.5))),
    ("c-x", lambda: node(SUB, const(0.5), var(0))),
    ("x^c", lambda: node(POW, var(0), const(1.5))),
    ("c^x", lambda: node(POW, const(1.5), var(0))),
    ("(c x)^(c x)", lambda: node(POW, node(MUL, const(0.5), var(0)), node(MUL, const(0.3), var(1)))),
    ("neg chain", lambda: node(NEG, node(MUL, node(NEG, node(ADD, var(0), const(0.2))), const(3.0)))),
    ("sin(x)*c + c", lambda: node(ADD, node(MUL, node(SIN, var(0)), const(1.1)), const(0.2))),
    ("(x+c)*(x-c)", lambda: node(MUL, node(ADD, var(0), const(0.3)), node(SUB, var(1), const(0.2)))),
    ("(c x)*(sin x)", lambda: node(MUL, node(MUL, const(0.3), var(0)), node(SIN, var(1)))),
    ("x*(c x)", lambda: node(MUL, var(2), node(MUL, const(0.3), var(0)))),
    ("(c + x) + (x + c)", lambda: node(ADD, node(ADD, const(0.3), var(0)), node(ADD, var(1), const(0.2)))),
    ("x - (c x)", lambda: node(SUB, var(2), node(MUL, const(0.3), var(0)))),
    ("deep sum", lambda: node(ADD, node(ADD, node(ADD, node(MUL, const(1.0), var(0)),
                                                    node(MUL, const(2.0), var(1))),
                                           node(MUL, const(3.0), var(2))),
                               node(MUL, node(EXP, node(MUL, const(0.2), var(0))), const(0.5)))),
]


@pytest.mark.parametrize("name,build", EDGE, ids=[e[0] for e in EDGE])
def test_edge_programs(name, build, X):
    assert check_tree(build(), X) == 0


def test_random_full_operator_trees(X):
    rng = np.random.default_rng(2024)
    done = rejected = 0
    while done < 250:
        t = full_tree(rng, int(rng.integers(1, 6)))
        if not t.constants() or host.gpu_fit_plan(t) is None:
            continue
        rc = check_tree(t, X, rel=1e-8)
        assert rc in (0, -3), rc
        rejected += rc == -3
        done += 1
    assert rejected <= 10, rejected


def test_structured_candidates_all_compile(X):
    rng = workloads.rng_for(3)
    for k in list(range(1, 21)) * 4:
        t = workloads.structured_candidate(rng, 3, k)
        assert check_tree(t, X, rel=1e-8) == 0, str(t)


def test_rejections():
    # An input beyond the VM's eight registers and an unsupported operator.
    rc, *_ = compile_tree(node(MUL, var(9), const(2.0)), nvars=10)
    assert rc == -3
    L = lib.load()
    info = np.zeros(6, np.int32)
    buf = np.zeros(64, np.uint32)
    mod = np.asarray([0, 0, 0, 0, 26, 0, 0, 0], np.uint32)  # an operator without an adjoint
    assert L.rch_fitvm_compile(lib.ptr(mod, lib.u32p), 8, 1, 3, COL_BYTES, 0, lib.ptr(buf, lib.u32p),
                               64, lib.ptr(buf, lib.u32p), 64, lib.ptr(info, lib.i32p)) == -2
    bad = np.asarray([0, 0, 0, 0, 2, 0, 0, 0], np.uint32)  # ADD with one operand: malformed
    assert L.rch_fitvm_compile(lib.ptr(bad, lib.u32p), 8, 1, 3, COL_BYTES, 0, lib.ptr(buf, lib.u32p),
                               64, lib.ptr(buf, lib.u32p), 64, lib.ptr(info, lib.i32p)) == -1
    small = np.zeros(2, np.uint32)
    rc, plan, *_ = compile_tree(node(MUL, var(0), const(2.0)))
    rec = np.ascontiguousarray(plan.words, np.uint32)
    assert L.rch_fitvm_compile(lib.ptr(rec, lib.u32p), len(rec), 1, 3, COL_BYTES, 0,
                               lib.ptr(small, lib.u32p), 0, lib.ptr(buf, lib.u32p), 64,
                               lib.ptr(info, lib.i32p)) == -4


def test_simulator_catches_a_planted_slot_error(X):
    # Plant a wrong read slot in the first add of a compiled program (a word
    # with the fused ADD, ACC): the simulator must refuse it (read of an
    # unwritten slot) or the values move.
    t = node(ADD, node(MUL, node(SIN, var(0)), const(1.1)), node(MUL, node(COS, var(1)), const(0.4)))
    rc, plan, fw, rv, info = compile_tree(t)
    assert rc == 0
    fw = fw.copy().reshape(-1, 2)
    add = [i for i in range(len(fw))
           if fw[i, 0] & (fitvm_sim.D["VMF_ADD"] | fitvm_sim.D["VMF_ACC"])][0]
    good, _ = fitvm_sim.run(fw.ravel(), rv, plan.c0, X, COL_BYTES, 0)
    ur = (int(fw[add, 1]) >> 12) & 15
    # another written slot: the value moves
    wrong = fw.copy()
    wrong[add, 1] = (int(wrong[add, 1]) & ~(15 << 12)) | (((ur + 1) % 3) << 12)
    moved, _ = fitvm_sim.run(wrong.ravel(), rv, plan.c0, X, COL_BYTES, 0)
    assert not np.allclose(moved, good)
    # a slot nothing wrote: refused
    wrong[add, 1] = (int(wrong[add, 1]) & ~(15 << 12)) | (14 << 12)
    with pytest.raises(fitvm_sim.SlotError):
        fitvm_sim.run(wrong.ravel(), rv, plan.c0, X, COL_BYTES, 0)


def test_simulator_catches_a_live_slot_named_instead_of_the_sink(X):
    # A tape word that keeps no partial names the B sink.  The simulator
    # writes NaN to the slot such a word names: plant the live slot of the
    # sin partial into the cos word that keeps nothing, and the sin term's
    # inner-constant gradient must come out wrong.
    t = node(ADD, node(MUL, node(SIN, node(MUL, const(0.7), var(0))), const(1.3)),
             node(MUL, node(COS, var(1)), const(0.4)))
    rc, plan, fw, rv, info = compile_tree(t, 3, INPUTS_IN_U)
    assert rc == 0
    fw = fw.copy().reshape(-1, 2)
    D = fitvm_sim.D
    unmc = [i for i in range(len(fw)) if fw[i, 0] & D["VMF_UNMC"]]
    need = [i for i in unmc if fw[i, 0] & D["VMF_NEED"]]
    keep_nothing = [i for i in unmc if not fw[i, 0] & D["VMF_NEED"]]
    assert len(need) == 1 and len(keep_nothing) == 1
    live = (int(fw[need[0], 1]) >> 16) & 15
    assert ((int(fw[keep_nothing[0], 1]) >> 16) & 15) == D["VM_B_SINK"]
    _, good = fitvm_sim.run(fw.ravel(), rv, plan.c0, X, COL_BYTES, INPUTS_IN_U)
    wrong = fw.copy()
    wrong[keep_nothing[0], 1] = (int(wrong[keep_nothing[0], 1]) & ~(15 << 16)) | (live << 16)
    _, bad = fitvm_sim.run(wrong.ravel(), rv, plan.c0, X, COL_BYTES, INPUTS_IN_U)
    assert any(not np.array_equal(bad[j], good[j], equal_nan=True) for j in good)


@pytest.mark.parametrize("flags", [INPUTS_IN_U, INPUTS_IN_U | VALUE_ONLY | SMALL_U])
def test_inputs_in_u_and_value_only_streams(flags, X):
    # The value-pass stream keeps the inputs in U slots 0..v-1 and drops the
    # tape; with an 8-slot file every slot index stays below 8.
    rng = np.random.default_rng(99)
    done = 0
    while done < 150:
        t = full_tree(rng, int(rng.integers(1, 6)))
        if not t.constants() or host.gpu_fit_plan(t) is None:
            continue
        rc = check_tree(t, X, rel=1e-8, flags=flags)
        assert rc in (0, -3), rc
        done += 1
    wl_rng = workloads.rng_for(3)
    for k in list(range(1, 21)) * 2:
        assert check_tree(workloads.structured_candidate(wl_rng, 3, k), X, rel=1e-8,
                          flags=flags) == 0
    for name, build in EDGE:
        assert check_tree(build(), X, flags=flags) == 0, name
