"""GPU parity of the exact order, rank and maximum stages
(rc_order_stat_score_batch, rc_rank_score_batch, rc_rank_debug_b,
rc_max_scan_batch) against the CPU oracle, through the C ABI.

Mirrors the reference's checks (gpu.rs:6375 rank sums bit-identical to the
CPU; metrics.rs:2191 ties): programs built from +, -, *, / evaluate to the
same f64 bits on both sides, so the median, the interquartile mean (summed
left to right on both sides), the maximum and its row, and the integer rank
sums must agree bit for bit; programs with transcendental operators are held
to f64 rounding.  Segments carry several constant vectors per program
(trials), ties come from rounded targets and constant predictions, -0.0 and
+0.0 predictions tie, and non-finite predictions are status 2.
"""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip import host
from rcsr_hip.expr import ADD, DIV, EXP, LOG, MUL, SIN, SUB, const, node, var

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


N = 3001


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(21)
    x = rng.uniform(-2.0, 2.0, size=(N, 2))
    x[:40, 0] = 0.0
    x[40:80, 0] = -0.0
    y = np.round(np.sin(x[:, 0]) + 0.3 * x[:, 1], 1)   # ties in the target
    ev = host.GpuEvaluator(np.ascontiguousarray(x.T, np.float32).reshape(-1), N, 2,
                           y.astype(np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(x.T).reshape(-1), y)
    yield x, y, ev
    ev.close()


def _exact_trees():
    return [
        node(ADD, node(MUL, const(1.3), var(0)), node(MUL, const(0.4), var(1))),
        node(DIV, node(SUB, var(0), const(0.2)), node(ADD, node(MUL, var(1), var(1)), const(1.5))),
        node(MUL, const(-1.0), var(0)),                    # -0.0 / +0.0 predictions tie
        node(ADD, node(MUL, const(0.0), var(0)), const(0.7)),  # constant: one tie run
    ]


def _smooth_trees():
    return [
        node(ADD, node(MUL, const(0.9), node(SIN, var(0))), node(MUL, const(0.3), var(1))),
        node(MUL, const(0.5), node(EXP, node(MUL, const(0.4), var(1)))),
    ]


def _plans_and_constants(trees, ntrials, seed):
    rng = np.random.default_rng(seed)
    plans = [host.gpu_fit_plan(t) for t in trees]
    assert all(p is not None for p in plans)
    kmax = max(p.nconst for p in plans)
    c = np.zeros((len(plans) * ntrials, kmax))
    for i, p in enumerate(plans):
        for t in range(ntrials):
            c[i * ntrials + t, :p.nconst] = np.asarray(p.c0) * (1.0 + 0.25 * t) + (0.1 * t if t else 0.0)
    return plans, c


def _oracle_preds(trees, plans, c, ntrials, x):
    out = []
    for i, (t, p) in enumerate(zip(trees, plans)):
        for k in range(ntrials):
            out.append(oracle.predict_flat(t.with_constants(c[i * ntrials + k, :p.nconst]), x))
    return out


@pytest.mark.parametrize("what", [0, 1])
def test_order_statistics_are_bit_identical_for_exact_programs(data, what):
    x, y, ev = data
    trees = _exact_trees()
    plans, c = _plans_and_constants(trees, 3, 1)
    val, st, ms = ev.order_stat_batch(plans, what, c, ntrials=3)
    assert ms >= 0
    preds = _oracle_preds(trees, plans, c, 3, x)
    f = oracle.median_abs if what == 0 else oracle.iqr_mean_abs
    for s, p in enumerate(preds):
        assert st[s] == 0
        want = f(p - y)
        assert val[s] == want, (s, val[s], want)
    _say(f"order stat {what}: {len(preds)} segments bit-identical, {ms:.2f} ms")


def test_order_statistics_of_transcendental_programs_and_status(data):
    x, y, ev = data
    trees = _smooth_trees() + [node(LOG, node(SUB, var(0), const(1.0)))]   # NaN rows
    plans, c = _plans_and_constants(trees, 2, 2)
    for what, f in ((0, oracle.median_abs), (1, oracle.iqr_mean_abs)):
        val, st, ms = ev.order_stat_batch(plans, what, c, ntrials=2)
        assert ms >= 0
        preds = _oracle_preds(trees, plans, c, 2, x)
        for s, p in enumerate(preds):
            if not np.all(np.isfinite(p)):
                assert st[s] == 2 and math.isnan(val[s])
                continue
            assert st[s] == 0
            assert val[s] == pytest.approx(f(p - y), rel=1e-13, abs=1e-15)


def test_target_ranks_and_rank_sums_are_bit_identical(data):
    x, y, ev = data
    b = ev.rank_debug_b()
    want_b = oracle.centred_ranks_doubled(y)
    assert np.array_equal(b, want_b)
    trees = _exact_trees()
    plans, c = _plans_and_constants(trees, 3, 3)
    sums, bb, st, ms = ev.rank_score_batch(plans, c, ntrials=3)
    assert ms >= 0
    _, _, want_bb = oracle.rank_sums(want_b, want_b)
    assert bb == want_bb
    preds = _oracle_preds(trees, plans, c, 3, x)
    for s, p in enumerate(preds):
        a = oracle.centred_ranks_doubled(p)
        ab, aa, _ = oracle.rank_sums(a, want_b)
        assert st[s] == 0 and sums[s] == (ab, aa), (s, sums[s], (ab, aa))
    # the constant program (its first trial) is one tie run: A = 0 everywhere
    assert sums[9] == (0, 0)
    _say(f"rank sums: {len(preds)} segments bit-identical, {ms:.2f} ms")


def test_rank_scores_of_transcendental_programs_and_non_finite(data):
    x, y, ev = data
    trees = _smooth_trees() + [node(LOG, node(SUB, var(0), const(1.0)))]
    plans, c = _plans_and_constants(trees, 2, 4)
    sums, bb, st, ms = ev.rank_score_batch(plans, c, ntrials=2)
    assert ms >= 0
    want_b = oracle.centred_ranks_doubled(y)
    preds = _oracle_preds(trees, plans, c, 2, x)
    for s, p in enumerate(preds):
        if not np.all(np.isfinite(p)):
            assert st[s] == 2
            continue
        a = oracle.centred_ranks_doubled(p)
        ab, aa, _ = oracle.rank_sums(a, want_b)
        rho_dev = sums[s][0] / math.sqrt(sums[s][1] * bb)
        rho_cpu = ab / math.sqrt(aa * bb)
        assert st[s] == 0 and rho_dev == pytest.approx(rho_cpu, abs=1e-9)


def test_max_scan_maximum_row_and_level_interval(data):
    x, y, ev = data
    trees = _exact_trees()[:2] + _smooth_trees()
    plans, c = _plans_and_constants(trees, 2, 5)
    outs, ms = ev.max_scan_batch(plans, c, ntrials=2)
    assert ms >= 0
    preds = _oracle_preds(trees, plans, c, 2, x)
    levels, coords = [], []
    for s, p in enumerate(preds):
        vm, row, lo, hi, stt = oracle.max_scan(p - y)
        o = outs[s]
        assert o.status == 0 and o.lo == -math.inf and o.hi == math.inf
        if s < 4:   # exact programs: the same bits and the same row
            assert (o.vmax, o.row) == (vm, row), s
        else:
            assert o.vmax == pytest.approx(vm, rel=1e-13)
        levels.append(0.9 * vm if s % 2 else 1.2 * vm)
        coords.append(0)
    outs, ms = ev.max_scan_batch(plans, c, ntrials=2, coords=coords, levels=levels)
    assert ms >= 0
    for s, p in enumerate(preds):
        i, k = divmod(s, 2)
        tc = trees[i].with_constants(c[s, :plans[i].nconst])
        b = np.empty(N)
        for r in range(N):
            gr, _ = oracle.gradients(tc, x[r])
            b[r] = gr[0]
        vm, row, lo, hi, stt = oracle.max_scan(p - y, b, levels[s])
        o = outs[s]
        assert o.status == 0
        if math.isinf(lo):
            assert o.lo == lo and o.hi == hi
        else:
            assert o.lo == pytest.approx(lo, rel=1e-12, abs=1e-12)
            assert o.hi == pytest.approx(hi, rel=1e-12, abs=1e-12)
    _say(f"max scan: {len(preds)} segments, {ms:.2f} ms")


def test_stage_arguments_are_validated(data):
    x, y, ev = data
    plans, c = _plans_and_constants(_exact_trees()[:1], 1, 6)
    val, st, ms = ev.order_stat_batch(plans, 2, c)      # unknown statistic
    assert ms < 0
    vals, st, ms = ev.order_stat_batch(plans, 0, c, ntrials=1)
    assert ms >= 0 and st[0] == 0
