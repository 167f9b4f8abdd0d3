"""CPU checks of the oracle's log-loss metric and logistic IRLS
(metrics.rs:1467-1475, 1547-1620; fit.rs:3271-3331), the CPU arm the device
logistic fit (metric 5) is compared with.

Pinned by the reference's own known answers (metrics.rs:1934-1943: log loss
at confident-correct, confident-wrong and zero logits; tests/logloss_gpu.rs:
305-327: a constant model reaches the closed form ln(n+/n-)), and by a
brute-force mean of ln(1 + exp(-t p)) in numpy.
"""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip.expr import ADD, MUL, const, node, var

LOG_LOSS = 10


def _affine(a, b):
    return node(ADD, node(MUL, const(a), var(0)), const(b))


def test_log_loss_known_answers():
    x = np.asarray([[-1.0], [1.0]])
    y = np.asarray([0.0, 1.0])
    big = node(MUL, const(1e6), var(0))          # logits -1e6, 1e6
    assert oracle.metric_score(big, x, y, metric=LOG_LOSS) == 0.0
    wrong = node(MUL, const(-1e6), var(0))       # logits 1e6, -1e6
    assert oracle.metric_score(wrong, x, y, metric=LOG_LOSS) == pytest.approx(1e6, rel=1e-6)
    zero = node(MUL, const(0.0), var(0))
    assert oracle.metric_score(zero, x, y, metric=LOG_LOSS) == pytest.approx(math.log(2.0), abs=1e-15)


def test_log_loss_is_the_mean_softplus_of_the_signed_margin():
    rng = np.random.default_rng(3)
    x = rng.uniform(-2, 2, size=(300, 1))
    y = (rng.uniform(size=300) < 0.4).astype(float) * 3.0 + 1.0   # labels by the midpoint 2.5
    t = np.where(y > 2.5, 1.0, -1.0)
    p = 0.7 * x[:, 0] - 0.2
    want = np.mean(np.logaddexp(0.0, -t * p))
    got = oracle.metric_score(_affine(0.7, -0.2), x, y, metric=LOG_LOSS)
    assert got == pytest.approx(want, rel=1e-13)


@pytest.mark.parametrize("positives", [0.5, 0.1, 0.01])
def test_constant_model_reaches_the_closed_form(positives):
    n = 1000
    k = round(n * positives)
    y = np.asarray([1.0 if i < k else 0.0 for i in range(n)])
    x = np.arange(n, dtype=float).reshape(-1, 1)
    score, c = oracle.fit_metric(const(0.0), x, y, iters=16, metric=LOG_LOSS)
    want = math.log(k / (n - k))
    assert abs(c[0] - want) / max(abs(want), 1.0) <= 1e-6
    # the reported value is the log loss of the fitted constant
    assert score == pytest.approx(oracle.metric_score(const(c[0]), x, y, metric=LOG_LOSS), rel=1e-14)


def test_affine_fit_improves_and_is_a_stationary_point():
    rng = np.random.default_rng(41)
    x = rng.uniform(-2, 2, size=(2000, 1))
    y = (rng.uniform(size=2000) < 1 / (1 + np.exp(-(1.5 * x[:, 0] - 0.3)))).astype(float)
    start = oracle.metric_score(_affine(0.0, 0.0), x, y, metric=LOG_LOSS)
    score, c = oracle.fit_metric(_affine(0.0, 0.0), x, y, iters=8, metric=LOG_LOSS)
    assert score < start
    # the gradient of the mean log loss vanishes at the optimum
    t = np.where(y > 0.5, 1.0, -1.0)
    p = c[0] * x[:, 0] + c[1]
    g = -(t / (1 + np.exp(t * p)))
    assert abs(np.mean(g * x[:, 0])) < 1e-7 and abs(np.mean(g)) < 1e-7
