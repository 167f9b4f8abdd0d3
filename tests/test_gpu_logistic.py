"""GPU checks of the logistic fit arm (rc_fit_batch metric 5; eval.cu:683-705,
1569-1589, 1661-1663, 1750-1757, 1831-1835) through the C ABI.

Mirrors the reference's device log-loss tests (tests/logloss_gpu.rs:480-560):
a constant model reaches the closed form ln(n+/n-), a two-parameter logistic
model is as good as the CPU logistic IRLS on the objective (rel 1e-6) with
constants within 1e-4, extreme start logits stay finite, and the reported
loss is the exact sum of ln(1 + exp(-t p)) at the returned constants.
"""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip import host
from rcsr_hip.expr import ADD, EXP, MUL, SIN, const, node, var

pytestmark = pytest.mark.gpu

LOGISTIC = 5
LOG_LOSS = 10


def _say(msg):
    print(msg, flush=True)


def _evaluator(x, y, signs=True):
    n, v = x.shape
    ev = host.GpuEvaluator(np.ascontiguousarray(x.T, np.float32).reshape(-1), n, v,
                           y.astype(np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(x.T).reshape(-1), y)
    if signs:
        lo, hi = y.min(), y.max()
        assert ev.set_fit_sign(np.where(y > 0.5 * (lo + hi), 1.0, -1.0))
    return ev


def _softplus_sum(p, y):
    t = np.where(y > 0.5 * (y.min() + y.max()), 1.0, -1.0)
    return float(np.sum(np.logaddexp(0.0, -t * p)))


def _affine(a, b):
    return node(ADD, node(MUL, const(a), var(0)), const(b))


def _sample(n=2000, seed=41):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2, 2, size=(n, 1))
    y = (rng.uniform(size=n) < 1 / (1 + np.exp(-(1.5 * x[:, 0] - 0.3)))).astype(float)
    return x, y


@pytest.mark.parametrize("positives", [0.5, 0.2, 0.02])
def test_constant_model_matches_the_closed_form(positives):
    n = 600
    k = round(n * positives)
    x = (np.arange(n, dtype=float) / n).reshape(-1, 1)
    y = np.asarray([1.0 if i < k else 0.0 for i in range(n)])
    _say(f"constant model, positives {positives}")
    ev = _evaluator(x, y)
    try:
        outs, ms = ev.fit_batch([host.gpu_fit_plan(const(0.0))], 8, LOGISTIC)
    finally:
        ev.close()
    assert ms >= 0 and host.fit_status_usable(outs[0].status), outs[0].status
    want = math.log(k / (n - k))
    got = outs[0].c[0]
    assert abs(got - want) / max(abs(want), 1.0) <= 1e-6, (got, want)
    assert outs[0].sse == pytest.approx(_softplus_sum(np.full(n, got), y), rel=1e-12)


def test_affine_parity_with_cpu_logistic_irls():
    x, y = _sample()
    start = _affine(0.0, 0.0)
    start_score = oracle.metric_score(start, x, y, metric=LOG_LOSS)
    cpu_score, cpu_c = oracle.fit_metric(start, x, y, iters=8, metric=LOG_LOSS)
    ev = _evaluator(x, y)
    try:
        outs, ms = ev.fit_batch([host.gpu_fit_plan(start)], 8, LOGISTIC)
    finally:
        ev.close()
    o = outs[0]
    assert ms >= 0 and host.fit_status_usable(o.status), o.status
    dev_score = o.sse / len(y)
    assert dev_score == pytest.approx(
        oracle.metric_score(start.with_constants(o.c), x, y, metric=LOG_LOSS), rel=1e-12)
    assert dev_score <= start_score
    assert abs(dev_score - cpu_score) / cpu_score <= 1e-6, (dev_score, cpu_score)
    for a, b in zip(o.c, cpu_c):
        assert abs(a - b) / max(abs(b), 1.0) <= 1e-4, (o.c, cpu_c)
    _say(f"affine: device {dev_score:.12f} cpu {cpu_score:.12f}")


def test_extreme_start_logits_stay_finite():
    x, y = _sample(400, 42)
    starts = (300.0, -300.0, 5000.0)
    ev = _evaluator(x, y)
    try:
        outs, ms = ev.fit_batch([host.gpu_fit_plan(_affine(a, 0.0)) for a in starts], 8, LOGISTIC)
    finally:
        ev.close()
    assert ms >= 0
    for a, o in zip(starts, outs):
        assert np.all(np.isfinite(o.c)) and math.isfinite(o.sse) and o.sse >= 0.0
        before = _softplus_sum(a * x[:, 0], y)
        # summation order may differ from numpy's in the last place
        assert o.sse <= before * (1 + 1e-12), (a, o)
        assert o.sse == pytest.approx(
            _softplus_sum(oracle.predict(_affine(a, 0.0).with_constants(o.c), x), y), rel=1e-12)


def test_alone_or_in_a_batch_gives_the_same_bits():
    x, y = _sample(1500, 5)
    t = node(ADD, node(MUL, const(0.5), node(SIN, node(MUL, const(1.1), var(0)))), const(0.1))
    ev = _evaluator(x, y)
    try:
        alone, ms0 = ev.fit_batch([host.gpu_fit_plan(t)], 12, LOGISTIC)
        again, ms1 = ev.fit_batch([host.gpu_fit_plan(t)], 12, LOGISTIC)
        many = [host.gpu_fit_plan(_affine(0.1 * i, -0.2)) for i in range(40)]
        many.insert(17, host.gpu_fit_plan(t))
        batch, ms2 = ev.fit_batch(many, 12, LOGISTIC)
    finally:
        ev.close()
    assert min(ms0, ms1, ms2) >= 0
    for o in (again[0], batch[17]):
        assert o.sse == alone[0].sse and np.array_equal(o.c, alone[0].c) and o.status == alone[0].status


def test_a_batch_reports_exact_losses_and_never_worsens():
    x, y = _sample(3000, 9)
    x = np.concatenate([x, np.random.default_rng(2).uniform(-1, 1, size=(len(y), 1))], 1)
    trees = [
        _affine(0.1, 0.0),
        node(ADD, node(MUL, const(0.5), node(SIN, node(MUL, const(1.1), var(0)))), const(0.1)),
        node(ADD, node(MUL, const(0.2), var(1)), node(MUL, const(0.3), var(0))),
        node(MUL, const(0.4), node(EXP, node(MUL, const(0.3), var(0)))),
    ] * 16
    ev = _evaluator(x, y)
    try:
        plans = [host.gpu_fit_plan(t) for t in trees]
        outs, ms = ev.fit_batch(plans, 20, LOGISTIC)
        sq, ms2 = ev.fit_batch(plans, 20, 0)    # the squared arm still fits the same batch
    finally:
        ev.close()
    assert ms >= 0 and ms2 >= 0
    for t, o in zip(trees, outs):
        assert host.fit_status_usable(o.status), o.status
        start = _softplus_sum(oracle.predict(t, x), y)
        got = _softplus_sum(oracle.predict(t.with_constants(o.c), x), y)
        assert o.sse <= start * (1 + 1e-12)
        assert o.sse == pytest.approx(got, rel=1e-10)
    # identical candidates give identical results
    assert all(outs[i].sse == outs[i % 4].sse for i in range(len(outs)))
    assert all(math.isfinite(o.sse) for o in sq)
    _say(f"batch of {len(trees)}: {ms:.2f} ms")


def test_the_launch_is_refused_without_label_signs():
    x, y = _sample(200, 3)
    ev = _evaluator(x, y, signs=False)
    try:
        outs, ms = ev.fit_batch([host.gpu_fit_plan(_affine(0.0, 0.0))], 8, LOGISTIC)
        assert ms < 0
        outs, ms = ev.fit_batch([host.gpu_fit_plan(_affine(0.0, 0.0))], 8, 0)
        assert ms >= 0
    finally:
        ev.close()


def test_plain_fp64_arm_and_mixed_batches_fit_the_same_objective():
    x, y = _sample(1200, 11)
    start = _affine(0.2, 0.1)
    ev = _evaluator(x, y)
    try:
        mixed, ms0 = ev.fit_batch([host.gpu_fit_plan(start)], 12, LOGISTIC)
        het, ms1 = ev.fit_batch_hetero([host.gpu_fit_plan(start)] * 3, [0, 0, 0],
                                       [LOGISTIC, 0, LOGISTIC], 12)
        assert ev.set_fit_mode(1) == 0          # returns the previous mode
        plain, ms2 = ev.fit_batch([host.gpu_fit_plan(start)], 12, LOGISTIC)
        assert ev.set_fit_mode(0) == 1
    finally:
        ev.close()
    assert min(ms0, ms1, ms2) >= 0
    # a candidate's result does not depend on its neighbours' metrics
    assert het[0].sse == mixed[0].sse and np.array_equal(het[0].c, mixed[0].c)
    assert het[2].sse == mixed[0].sse
    # the squared neighbour reports its own objective
    assert het[1].sse == pytest.approx(
        float(np.sum((oracle.predict(start.with_constants(het[1].c), x) - y) ** 2)), rel=1e-10)
    for o in (mixed[0], plain[0]):
        assert host.fit_status_usable(o.status), o.status
        assert o.sse == pytest.approx(
            _softplus_sum(oracle.predict(start.with_constants(o.c), x), y), rel=1e-12)
    assert abs(plain[0].sse - mixed[0].sse) / mixed[0].sse <= 1e-6
