"""GPU parity of the plain FP64 arm (rc_set_fit_mode 1) against the oracle.

The FP64 arm runs the mixed arm's decomposition (work items of one
candidate's row tile, the recorded launch sequence, the control kernel) with
every row pass on f64 rows: f64 value and adjoint sweeps, an FP64-MFMA Gram
and f64 trial losses.  Each assembly class is exercised: the fit VM at 16 and
32 Gram columns, and the general interpreter for a program with more inputs
than the VM's files hold and for an oversized right-deep program.
"""
import math
import sys

import numpy as np
import pytest

import oracle
from rcsr_hip import host, workloads
from rcsr_hip.expr import ADD, COS, MUL, SIN, const, node, var

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _evaluator(x, y, w=None):
    n, v = x.shape
    ev = host.GpuEvaluator(x.T.astype(np.float32).reshape(-1), n, v, np.asarray(y, np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(x.T).reshape(-1), np.asarray(y, np.float64), w)
    assert ev.set_fit_mode(host.FIT_MODE_FP64) == 0
    return ev


def _design(nv, n=1200, seed=21):
    t = 2 * np.pi * np.arange(n) / (n - 1)
    x = np.stack([np.sin((k + 1) * t + 0.1 * k) for k in range(nv)], 1)
    rng = np.random.default_rng(seed)
    return x, rng


def _left_sum(terms):
    tree = terms[0]
    for t in terms[1:]:
        tree = node(ADD, tree, t)
    return tree


def _right_sum(terms):
    tree = terms[-1]
    for t in reversed(terms[:-1]):
        tree = node(ADD, t, tree)
    return tree


def _check_against_oracle(trees, x, y, iters=8, tol=1e-10):
    ev = _evaluator(x, y)
    outs, ms = ev.fit_batch([host.gpu_fit_plan(t) for t in trees], iters, 0)
    ev.close()
    assert ms >= 0.0
    for t, o in zip(trees, outs):
        cpu_sse, _ = oracle.fit_weighted(t, x, y, None, iters)
        assert o.status in (0, 4), (str(t), o.status)
        got = oracle.sse(t.with_constants(o.c), x, y)
        assert abs(got - cpu_sse) <= tol * cpu_sse, (str(t), got, cpu_sse)
        assert o.sse == pytest.approx(got, rel=1e-12)


def test_fp64_arm_vm_classes_reach_the_cpu_optimum():
    # Linear sums of sin-columns with k = 3, 9, 14 (16 Gram columns) and
    # 17, 20 (32 Gram columns) run on the fit VM.
    x, rng = _design(5)
    y = x @ np.array([1.0, -0.5, 0.25, 2.0, -1.5]) + rng.normal(0, 1e-3, len(x))
    trees = []
    for k in (3, 9, 14, 17, 20):
        # c * x_j terms, every fourth one c * sin(x_j) (at most 96 nodes).
        terms = [node(MUL, const(0.3 + 0.05 * j), node(SIN, var(j % 5)))
                 if j % 4 == 3 else node(MUL, const(0.2 + 0.1 * j), var(j % 5)) for j in range(k)]
        trees.append(_left_sum(terms))
        assert host.gpu_fit_plan(trees[-1]) is not None
    _say(f"fp64 arm VM classes: k = 3, 9, 14, 17, 20 over n={len(x)}")
    _check_against_oracle(trees, x, y, tol=1e-9)


def test_fp64_arm_general_interpreter_classes_reach_the_cpu_optimum():
    # Nine inputs (more than the VM's input slots) with a left-deep sum, and
    # a right-deep sum of 12 terms whose stack outgrows the VM's files.
    x9, rng = _design(9)
    y9 = x9 @ np.linspace(-1.0, 1.5, 9) + rng.normal(0, 1e-3, len(x9))
    wide = _left_sum([node(MUL, const(0.5), var(j)) for j in range(9)])
    x3, rng = _design(3, seed=5)
    y3 = 1.2 * x3[:, 0] - 0.4 * np.cos(x3[:, 1]) + 0.7 * x3[:, 2] + rng.normal(0, 1e-3, len(x3))
    deep = _right_sum([node(MUL, const(0.1 * (j + 1)), var(j % 3)) if j % 3 != 1
                       else node(MUL, const(-0.2), node(COS, var(1))) for j in range(12)])
    _say("fp64 arm general interpreter: 9 inputs; right-deep 12 terms")
    _check_against_oracle([wide], x9, y9, tol=1e-9)
    _check_against_oracle([deep], x3, y3, tol=1e-9)


def test_fp64_arm_tracks_the_oracle_on_random_programs():
    # C1 programs (ops +,-,*,/,exp,log,sin,cos; 0-4 constants): both sides
    # run every row in f64, so the trajectories agree to summation order.
    wl = workloads.config1()
    x, y = np.asarray(wl.x), np.asarray(wl.y)
    trees = [t for t in wl.trees if host.gpu_fit_plan(t) is not None][:200]
    ev = _evaluator(x, y)
    outs, ms = ev.fit_batch([host.gpu_fit_plan(t) for t in trees], 8, 0)
    ev.close()
    assert ms >= 0.0
    cpu, _ = oracle.fit_batch(trees, x, y, None, 8, threads=8, kmax=20)
    rels, rels0 = [], []
    for t, o, cl in zip(trees, outs, cpu):
        if not host.fit_status_usable(o.status) or not math.isfinite(cl) or cl <= 0:
            continue
        got = oracle.sse(t.with_constants(o.c), x, y)
        assert o.sse == pytest.approx(got, rel=1e-9, abs=1e-300)
        rels.append(abs(got - cl) / cl)
        if o.status == 0:
            rels0.append(rels[-1])
    rels, rels0 = np.asarray(rels), np.asarray(rels0)
    _say(f"fp64 arm C1: {len(rels)} compared; median {np.median(rels):.2e}; "
         f"max {rels.max():.2e}; within 1e-9: {np.mean(rels <= 1e-9):.3f}; status 0: "
         f"{len(rels0)}, within 1e-9: {np.mean(rels0 <= 1e-9):.3f}")
    assert len(rels) >= 150
    assert np.median(rels) <= 1e-12
    # Both sides run every row in f64 but sum the Gram in different orders
    # (FP64 MFMA blocks against a sequential sum), which is enough to part
    # the accept / reject sequence of an ill-conditioned (status 4) or
    # unconverged candidate; well-conditioned fits agree.
    assert len(rels0) >= 20
    assert np.mean(rels0 <= 1e-9) >= 0.95
