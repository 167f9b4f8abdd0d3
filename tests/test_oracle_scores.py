"""Independent pins of the oracle's score reductions and of its log-squashed
IRLS arm.

The oracle's `or_metric_score` restates metrics.rs:985-1060
(`score_weighted`: a weighted mean Σ wᵢ·ρ(eᵢ) / Σ wᵢ, zero-weight rows
skipped before any arithmetic, a negative or non-finite weight or an all-zero
weight vector scoring +inf, a non-finite prediction on a positive-weight row
scoring +inf) and `or_implicit_score` restates implicit.rs:122-155.  Here the
predictions come from numpy expressions of the trees, never from the
oracle's evaluator, and the reductions are exact (`math.fsum`), so a
restatement that divides by n instead of Σw, forgets the mask, or squares
where it should take ln(1 + e²) fails.  The duplicate-row identities ("a
weight of 2 means this row twice", metrics.rs:975-978) hold for any correct
weighted mean and for no plausible wrong one.

The log-squashed branch of `irls_fit` (fit.rs:3092-3170 with the weight
1/(1 + r²), fit.rs:2972) is pinned by the reference's own invariant
(fit.rs:4500-4510: it beats the squared fit on its own objective), by its
reported loss being the numpy objective of its returned constants, by never
ending above its start, and by a scipy minimisation of the same objective
bounding it from below and, on a well-conditioned fixture, meeting it.
"""
import math

import numpy as np
import pytest
from scipy.optimize import minimize

import oracle
from rcsr_hip.expr import ADD, LOG, MUL, SIN, SQUARE, SUB, const, node, var

INF = math.inf


def _rho(metric, e):
    if metric == 0:
        return e * e
    if metric == 4:
        return np.abs(e)
    if metric == 1:
        return np.log(1.0 + e * e)
    raise ValueError(metric)


def _weighted_mean(metric, pred, y, w=None):
    """metrics.rs:985-1060 from its documented definition, exactly summed."""
    pred = np.asarray(pred, np.float64)
    y = np.asarray(y, np.float64)
    if w is None:
        if not np.all(np.isfinite(pred)):
            return INF
        return math.fsum(_rho(metric, pred - y)) / len(pred)
    w = np.asarray(w, np.float64)
    if not np.all(np.isfinite(w)) or np.any(w < 0):
        return INF
    total = math.fsum(w)
    if not total > 0:
        return INF
    pos = w > 0
    if not np.all(np.isfinite(pred[pos])):
        return INF
    return math.fsum(w[pos] * _rho(metric, pred[pos] - y[pos])) / total


def _fixture(n=257, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2.0, 3.0, n)
    y = 0.8 * np.sin(x) - 0.3 * x + rng.normal(0, 0.2, n)
    y[::17] += 4.0
    return x, y


def _tree(a, b):
    # a*sin(x) + b*x; numpy restates it as a * np.sin(x) + b * x.
    return node(ADD, node(MUL, const(a), node(SIN, var(0))), node(MUL, const(b), var(0)))


@pytest.mark.parametrize("metric", [0, 1, 4])
def test_metric_score_unweighted_is_the_exact_mean(metric):
    x, y = _fixture()
    a, b = 0.9, -0.25
    got = oracle.metric_score(_tree(a, b), x[:, None], y, None, metric)
    want = _weighted_mean(metric, a * np.sin(x) + b * x, y)
    assert got == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("metric", [0, 1, 4])
def test_metric_score_weighted_with_zero_weights_is_the_exact_weighted_mean(metric):
    x, y = _fixture()
    rng = np.random.default_rng(8)
    w = rng.uniform(0.0, 3.0, len(x))
    w[::5] = 0.0
    a, b = 1.1, -0.4
    got = oracle.metric_score(_tree(a, b), x[:, None], y, w, metric)
    want = _weighted_mean(metric, a * np.sin(x) + b * x, y, w)
    assert got == pytest.approx(want, rel=1e-13)
    # Dividing by the row count instead of the weight total is far off.
    wrong = want * math.fsum(w) / len(x)
    assert abs(got - wrong) > 1e-3 * abs(want)


@pytest.mark.parametrize("metric", [0, 1, 4])
def test_a_weight_of_two_is_the_row_twice(metric):
    x, y = _fixture(n=101)
    reps = np.array([(i % 3) for i in range(len(x))])      # weights 0, 1, 2
    xd = np.repeat(x, reps)
    yd = np.repeat(y, reps)
    t = _tree(0.7, 0.15)
    weighted = oracle.metric_score(t, x[:, None], y, reps.astype(np.float64), metric)
    duplicated = oracle.metric_score(t, xd[:, None], yd, None, metric)
    assert weighted == pytest.approx(duplicated, rel=1e-13)


@pytest.mark.parametrize("metric", [0, 1, 4])
def test_metric_score_weight_and_finiteness_rules(metric):
    x = np.array([-1.0, 0.5, 2.0, 3.0])
    y = np.array([0.1, 0.2, -0.3, 0.4])
    t = node(LOG, var(0))                 # NaN at the negative row
    X = x[:, None]
    # The zero-weight row's NaN prediction is skipped before any arithmetic.
    w = np.array([0.0, 1.0, 2.0, 0.5])
    pred = np.log(np.where(x > 0, x, np.nan))
    want = _weighted_mean(metric, pred, y, w)
    assert math.isfinite(want)
    assert oracle.metric_score(t, X, y, w, metric) == pytest.approx(want, rel=1e-14)
    # The same row with a positive weight, or no weights at all: +inf.
    assert oracle.metric_score(t, X, y, np.array([1e-9, 1.0, 2.0, 0.5]), metric) == INF
    assert oracle.metric_score(t, X, y, None, metric) == INF
    good = node(MUL, const(0.5), var(0))
    assert oracle.metric_score(good, X, y, np.array([1.0, -1.0, 1.0, 1.0]), metric) == INF
    assert oracle.metric_score(good, X, y, np.array([1.0, np.nan, 1.0, 1.0]), metric) == INF
    assert oracle.metric_score(good, X, y, np.array([1.0, np.inf, 1.0, 1.0]), metric) == INF
    assert oracle.metric_score(good, X, y, np.zeros(4), metric) == INF


def _circle():
    return node(SUB, node(ADD, node(SQUARE, var(0)), node(SQUARE, var(1))), const(1.0))


def _circle_rows(n=300, seed=4):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    r = 1.0 + rng.normal(0, 0.05, n)
    return np.stack([r * np.cos(th), r * np.sin(th)], 1)


def _implicit_mean(X, w, metric):
    """implicit.rs:122-155 from its definition: for x1² + x2² − 1 the
    distance |f|/‖∇f‖ is |r² − 1| / (2r) and the plain residual is |r² − 1|."""
    r = np.hypot(X[:, 0], X[:, 1])
    d = np.abs(r * r - 1.0) / (2.0 * r) if metric == 0 else np.abs(r * r - 1.0)
    if w is None:
        return math.fsum(d) / len(d)
    pos = w > 0
    return math.fsum(w[pos] * d[pos]) / math.fsum(w[pos])


@pytest.mark.parametrize("metric", [0, 1])
def test_implicit_score_is_the_exact_weighted_mean_distance(metric):
    X = _circle_rows()
    got = oracle.implicit_score(_circle(), X, None, metric)
    assert got == pytest.approx(_implicit_mean(X, None, metric), rel=1e-13)
    rng = np.random.default_rng(9)
    w = rng.uniform(0.0, 2.0, len(X))
    w[::7] = 0.0
    got = oracle.implicit_score(_circle(), X, w, metric)
    assert got == pytest.approx(_implicit_mean(X, w, metric), rel=1e-13)
    # A weight of 2 is the row twice.
    reps = np.array([(i % 3) for i in range(len(X))])
    dup = oracle.implicit_score(_circle(), np.repeat(X, reps, axis=0), None, metric)
    assert oracle.implicit_score(_circle(), X, reps.astype(np.float64), metric) == \
        pytest.approx(dup, rel=1e-13)


def test_implicit_score_row_rules():
    X = _circle_rows(n=20)
    X[3] = [0.0, 0.0]                      # zero gradient: the distance is undefined
    w = np.ones(len(X))
    assert oracle.implicit_score(_circle(), X, w, 0) == INF
    w[3] = 0.0                             # masked: skipped before any arithmetic
    keep = np.arange(len(X)) != 3
    assert oracle.implicit_score(_circle(), X, w, 0) == pytest.approx(
        _implicit_mean(X[keep], None, 0), rel=1e-13)
    # The plain residual is defined there (|f| = 1).
    assert oracle.implicit_score(_circle(), X, None, 1) == pytest.approx(
        _implicit_mean(X, None, 1), rel=1e-13)
    w2 = np.ones(len(X))
    w2[5] = -1.0
    assert oracle.implicit_score(_circle(), X, w2, 0) == INF
    assert oracle.implicit_score(_circle(), X, np.zeros(len(X)), 1) == INF


# ---- the log-squashed IRLS arm ------------------------------------------------

def _ref_outlier_fixture():
    """fit.rs:4476-4485."""
    xs = [-2.0, -1.5, -1.0, -0.5, 0.0, 0.5, 1.0, 1.5, 2.0]
    y = [20.0 if i == len(xs) - 1 else 2.0 * v + 1.0 for i, v in enumerate(xs)]
    return np.asarray(xs)[:, None], np.asarray(y)


def _line(a, b):
    """fit.rs:4436-4441: a*x + b."""
    return node(ADD, node(MUL, const(a), var(0)), const(b))


def _log_squashed_line(c, x, y):
    e = c[0] * x + c[1] - y
    return math.fsum(np.log(1.0 + e * e)) / len(e)


def test_log_squashed_fit_beats_squared_on_its_objective():
    # fit.rs:4500-4510.
    x, y = _ref_outlier_fixture()
    loss, c = oracle.fit_metric(_line(0.5, 0.0), x, y, None, 16, 1)
    _, c2 = oracle.fit_metric(_line(0.5, 0.0), x, y, None, 16, 0)
    a = oracle.metric_score(_line(*c), x, y, None, 1)
    b = oracle.metric_score(_line(*c2), x, y, None, 1)
    assert a < b
    # The reported loss is the numpy objective of the returned constants.
    assert loss == pytest.approx(_log_squashed_line(c, x[:, 0], y), rel=1e-13)


def _robust_fixture(n=400):
    """gpu.rs:5277-5289."""
    t = np.arange(n) / (n - 1) * 4.0 - 2.0
    noise = (((np.arange(n) * 37) % 11) - 5.0) * 1e-3
    y = 0.5 + 1.75 * t + noise + np.where(np.arange(n) % 8 == 3, 25.0, 0.0)
    return t, y


def test_log_squashed_irls_meets_the_scipy_minimum():
    t, y = _robust_fixture()
    start = [1.0, 0.2]                    # (slope, intercept) of a*x + b
    f = lambda c: _log_squashed_line(c, t, y)
    g = lambda c: np.array([
        math.fsum(2 * (c[0] * t + c[1] - y) / (1 + (c[0] * t + c[1] - y) ** 2) * t),
        math.fsum(2 * (c[0] * t + c[1] - y) / (1 + (c[0] * t + c[1] - y) ** 2))]) / len(t)
    best = minimize(f, start, jac=g, method="BFGS", options={"gtol": 1e-13, "maxiter": 2000})
    for iters in (8, 32):
        loss, c = oracle.fit_metric(_line(*start), t[:, None], y, None, iters, 1)
        assert loss == pytest.approx(f(c), rel=1e-13)
        assert loss <= f(start)
        # No fit lands below the objective's minimum ...
        assert loss >= best.fun * (1 - 1e-12)
        # ... and the IRLS majoriser reaches it on this fixture: a wrong
        # weight (1/|r|, 1/(1+|r|), 2/(1+r²) without the base weight) stalls
        # at another point.
        assert loss <= best.fun * (1 + 1e-6), (iters, loss, best.fun)
        assert np.max(np.abs(np.asarray(c) - best.x)) < 1e-3
    # The least-squares optimum is measurably worse on this objective.
    _, c2 = oracle.fit_metric(_line(*start), t[:, None], y, None, 8, 0)
    assert f(c2) > best.fun + 1e-3


def test_log_squashed_irls_honours_the_base_weight():
    # gpu.rs:5426-5488: two groups want slopes +1.75 and -1; the second has
    # weight 1e-2, and every tenth row is masked with a wild target.
    n = 400
    t = np.arange(n) / (n - 1) * 4.0 - 2.0
    w = np.ones(n)
    y = np.empty(n)
    for i in range(n):
        if i % 10 == 7:
            w[i] = 0.0
            y[i] = 50.0 - 3.0 * t[i]
        elif i % 2 == 0:
            y[i] = 1.75 * t[i]
        else:
            w[i] = 1e-2
            y[i] = -t[i]
    start = node(ADD, const(0.0), node(MUL, const(0.2), var(0)))
    loss, c = oracle.fit_metric(start, t[:, None], y, w, 8, 1)
    pred = c[0] + c[1] * t
    assert loss == pytest.approx(_weighted_mean(1, pred, y, w), rel=1e-13)
    assert abs(c[1] - 1.75) < 0.05
    _, cu = oracle.fit_metric(start, t[:, None], y, None, 8, 1)
    wrong = _weighted_mean(1, cu[0] + cu[1] * t, y, w)
    assert wrong > loss + 1e-2
