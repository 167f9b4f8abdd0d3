"""Pins the parts of the CPU oracle that the recorded bits do not cover.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed operand) fails it: derivatives against central
differences and closed forms, IRLS against the weighted-median closed form,
the implicit distance against the exact unit-circle distance, and the
invariants the reference tests hold (fit.rs:4088-4135, 4488-4628, 4838-4935;
tests/implicit.rs:144-235).
"""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip.expr import (ABS, ADD, COS, CUBE, DIV, ERF, ERFC, EXP, GAUSS, LOG, MUL, NEG, POW,
                           SIG, SIN, SINH, COSH, SQRT, SQUARE, SUB, TAN, TANH, ATAN2, MOD,
                           const, node, var)


def _fd_check(build, c, x, tol=1e-6):
    tree = build(c)
    g, _ = oracle.gradients(tree, x)
    h = 1e-6
    for k in range(len(c)):
        cp = list(c)
        cm = list(c)
        cp[k] += h
        cm[k] -= h
        fd = (oracle.eval_row(build(cp), x) - oracle.eval_row(build(cm), x)) / (2 * h)
        assert abs(g[k] - fd) <= 1e-4 * (1 + abs(fd)), (k, g[k], fd)


UNARY = [SIN, COS, TAN, EXP, LOG, SQRT, ABS, TANH, SIG, SQUARE, CUBE, NEG, GAUSS, SINH, COSH,
         ERF, ERFC]


@pytest.mark.parametrize("op", UNARY)
def test_adjoint_matches_central_differences_per_operator(op):
    # d/dc op(c*x + 0.3) at a point inside every operator's domain.
    build = lambda c: node(op, node(ADD, node(MUL, const(c[0]), var(0)), const(c[1])))
    _fd_check(build, [0.7, 0.3], [0.9])


def test_adjoint_binary_operators_and_reference_cases():
    _fd_check(lambda c: node(DIV, node(MUL, const(c[0]), var(0)), node(ADD, var(0), const(c[1]))),
              [1.3, 2.0], [0.4])
    _fd_check(lambda c: node(POW, node(ADD, var(0), const(c[0])), const(c[1])), [1.5, 2.3], [0.4])
    _fd_check(lambda c: node(ATAN2, const(c[0]), var(0)), [1.2], [2.0])
    _fd_check(lambda c: node(MOD, node(MUL, const(c[0]), var(0)), const(c[1])), [3.0, 4.0], [1.7])
    _fd_check(lambda c: node(SUB, node(MUL, const(c[0]), node(SIN, var(0))), const(c[1])),
              [2.0, -0.5], [0.3])


def test_adjoint_closed_forms_pin_signs_and_indices():
    # f = c0*x^2 - c1/x at x=2: df/dc0 = 4, df/dc1 = -0.5 exactly.
    t = node(SUB, node(MUL, const(1.5), node(SQUARE, var(0))), node(DIV, const(0.7), var(0)))
    g, gx = oracle.gradients(t, [2.0])
    assert g.tolist() == [4.0, -0.5]
    # df/dx = 2*c0*x + c1/x^2 = 6 + 0.175
    assert gx[0] == pytest.approx(6.175, rel=1e-15)
    # input gradient of x0*x1 + exp(x1) at (2, 0): (0, 2 + 1)
    t2 = node(ADD, node(MUL, var(0), var(1)), node(EXP, var(1)))
    _, gx2 = oracle.gradients(t2, [2.0, 0.0])
    assert gx2.tolist() == [0.0, 3.0]


@pytest.mark.parametrize("x", [-4.0, -1.2, -0.3, 0.0, 1e-8, 0.2, 0.9, 1.5, 2.7, 5.5])
def test_erf_matches_the_c_library(x):
    assert oracle.erf(x) == pytest.approx(math.erf(x), rel=1e-13, abs=1e-16)
    assert oracle.erfc(x) == pytest.approx(math.erfc(x), rel=1e-12, abs=1e-300)


def test_eval_domain_guards_match_expr_eval():
    nan = oracle.eval_row(node(LOG, var(0)), [-1.0])
    assert math.isnan(nan)
    assert math.isnan(oracle.eval_row(node(DIV, const(1.0), var(0)), [1e-13]))
    assert math.isnan(oracle.eval_row(node(SQRT, var(0)), [-0.5]))
    assert math.isnan(oracle.eval_row(node(POW, var(0), const(0.5)), [-2.0]))
    assert oracle.eval_row(node(POW, var(0), const(2.0)), [-2.0]) == 4.0
    assert oracle.eval_row(node(EXP, var(0)), [701.0]) == math.inf
    assert oracle.eval_row(node(MOD, const(-7.0), const(3.0)), [0.0]) == 2.0


def _outlier_fixture():
    xs = [i / 10.0 for i in range(21)]
    x = [[v] for v in xs]
    y = [20.0 if i == len(xs) - 1 else 2.0 * v + 1.0 for i, v in enumerate(xs)]
    return x, y


def _line(a, b):
    return node(ADD, node(MUL, const(a), var(0)), const(b))


def test_l1_fit_beats_squared_on_l1_and_squared_wins_its_own():
    x, y = _outlier_fixture()
    l1_loss, c1 = oracle.fit_metric(_line(0.5, 0.0), x, y, None, 16, 4)
    _, c2 = oracle.fit_metric(_line(0.5, 0.0), x, y, None, 16, 0)
    t1 = _line(*c1)
    t2 = _line(*c2)
    assert oracle.metric_score(t1, x, y, None, 4) <= oracle.metric_score(t2, x, y, None, 4)
    assert abs(l1_loss - oracle.metric_score(t1, x, y, None, 4)) < 1e-12
    assert oracle.metric_score(t2, x, y, None, 0) < oracle.metric_score(t1, x, y, None, 0)


def test_l1_fit_recovers_a_clean_proportional_law():
    x = [[-1.0 + 2.0 * i / 20.0] for i in range(21)]
    y = [r[0] for r in x]
    loss, _ = oracle.fit_metric(node(MUL, const(3.0), var(0)), x, y, None, 8, 4)
    assert loss < 1e-12


def test_l1_fit_matches_the_weighted_median_closed_form():
    ratios = [2.0, 2.1, 1.9, 2.05, 1.95, 30.0, 2.02, 1.98, -12.0, 2.03, 1.97]
    xs = [0.5, 0.7, 0.9, 1.1, 1.3, 1.5, 1.7, 1.9, 2.1, 2.3, 2.5]
    base = [1.0, 1.0, 1.0, 1.0, 1.0, 0.5, 1.0, 1.0, 0.75, 1.0, 1.0]
    x = [[v] for v in xs]
    y = [a * r for a, r in zip(xs, ratios)]
    order = sorted(range(len(xs)), key=lambda i: ratios[i])
    total = sum(base[i] * xs[i] for i in order)
    acc, median = 0.0, None
    for i in order:
        acc += base[i] * xs[i]
        if acc >= 0.5 * total:
            median = ratios[i]
            break
    loss, c = oracle.fit_metric(node(MUL, const(1.0), var(0)), x, y, base, 64, 4)
    assert abs(c[0] - median) <= 1e-2 * max(abs(median), 1.0)
    at_median = oracle.metric_score(node(MUL, const(median), var(0)), x, y, base, 4)
    assert loss <= at_median * (1 + 1e-3)


def test_zero_iterations_is_a_no_op_for_every_metric():
    x = [[-2.0 + 4.0 * i / 40.0] for i in range(41)]
    y = [1.7 * r[0] ** 2 - 0.4 * r[0] + 0.8 for r in x]
    t = _line(3.0, 0.5)
    for metric in (0, 1, 4):
        score, c = oracle.fit_metric(t, x, y, None, 0, metric)
        assert c.tolist() == [3.0, 0.5]
        assert math.isfinite(score)
    s0, _ = oracle.fit_weighted(t, x, y, None, 0)
    assert s0 == oracle.sse(t, x, y)


def test_unit_circle_distance_is_exact():
    f = node(SUB, node(ADD, node(SQUARE, var(0)), node(SQUARE, var(1))), const(1.0))
    for (a, b) in [(1.2, 0.3), (0.5, -0.4), (-0.9, 0.9), (0.0, 1.5)]:
        r = math.hypot(a, b)
        d = oracle.signed_distance(f, [a, b])
        assert d == pytest.approx((r * r - 1) / (2 * r), rel=1e-14)
    # Scale-free: distance of 3f equals that of f.
    f3 = node(MUL, const(3.0), f)
    assert oracle.signed_distance(f3, [1.2, 0.3]) == pytest.approx(
        oracle.signed_distance(f, [1.2, 0.3]), rel=1e-15)
    # Zero-gradient guard: a constant candidate has no level set.
    assert oracle.signed_distance(node(ADD, const(1.0), const(0.0)), [0.3, 0.2]) is None


def test_implicit_fit_recovers_the_unit_circle():
    rng = np.random.default_rng(5)
    th = rng.uniform(0, 2 * np.pi, 200)
    X = np.stack([np.cos(th), np.sin(th)], 1)
    f = node(SUB, node(ADD, node(SQUARE, var(0)), node(SQUARE, var(1))), const(0.6))
    s, c = oracle.implicit_fit(f, X, None, 20)
    assert c[0] == pytest.approx(1.0, abs=1e-6)
    assert s < 1e-10


def test_linear_lm_reaches_the_least_squares_optimum():
    # gpu.rs:4388-4411 fixture: orthogonal sin((k+1)t) columns, p = 1..20.
    # For a linear model the LM optimum is the least-squares solution, which
    # numpy's SVD solver gives in closed form: a dropped damping term, a
    # transposed Gram or a wrong step sign leaves the SSE above it.
    n = 800
    for p in range(1, 21):
        t = 2 * np.pi * np.arange(n) / (n - 1)
        x = np.stack([np.sin((k + 1) * t) for k in range(p)], 1)
        noise = (((np.arange(n) * 37) % 11) - 5.0) * 2e-4
        y = x @ np.asarray([1.0 + 0.31 * k for k in range(p)]) + noise
        terms = [node(MUL, const(0.25 + 0.07 * k), var(k)) for k in range(p)]
        tree = terms[-1]
        for term in reversed(terms[:-1]):
            tree = node(ADD, term, tree)
        sse, c = oracle.fit_weighted(tree, x, y, None, 8)
        coef, *_ = np.linalg.lstsq(x, y, rcond=None)
        best = float(np.sum((x @ coef - y) ** 2))
        assert abs(sse - best) <= 1e-9 * best, (p, sse, best)
        assert np.max(np.abs(np.asarray(c) - coef)) <= 1e-6, p
