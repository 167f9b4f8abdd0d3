"""Launch shapes (rc_set_tuning) and the autotuner on the device, through the
C ABI: every shape gives the oracle's fits and the same per-row predictions,
and the autotuner times, applies and caches a shape."""
import math

import numpy as np
import pytest

import oracle
from rcsr_hip import autotune as at
from rcsr_hip import host, workloads
from rcsr_hip.expr import ADD, MUL, const, node, var
from rcsr_hip.host import TUNE_EVAL_TILE_ROWS, TUNE_FIT_ROWS_PER_LANE, TUNE_FIT_TILE_ROWS

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _evaluator(x, y):
    n, v = x.shape
    ev = host.GpuEvaluator(np.ascontiguousarray(x.T, np.float32).reshape(-1), n, v,
                           np.asarray(y, np.float32))
    assert ev.set_fit_data(np.ascontiguousarray(x.T).reshape(-1), np.asarray(y, np.float64))
    return ev


def test_tuning_entry_points_validate():
    x = np.random.default_rng(1).uniform(-1, 1, (1000, 2))
    ev = _evaluator(x, x[:, 0])
    assert ev.get_tuning(TUNE_FIT_TILE_ROWS) == 0
    assert ev.set_tuning(TUNE_FIT_TILE_ROWS, 4096) == 0
    assert ev.get_tuning(TUNE_FIT_TILE_ROWS) == 4096
    assert ev.set_tuning(TUNE_FIT_TILE_ROWS, 300) == -1       # not a multiple of 256
    assert ev.set_tuning(TUNE_FIT_TILE_ROWS, 1 << 23) == -1   # too large
    assert ev.set_tuning(TUNE_FIT_ROWS_PER_LANE, 3) == -1
    assert ev.set_tuning(TUNE_FIT_ROWS_PER_LANE, 2) == 0
    assert ev.set_tuning(TUNE_EVAL_TILE_ROWS, 1024) == 0
    assert ev.set_tuning(99, 0) == -1
    assert ev.get_tuning(99) == -1
    assert ev.set_tuning(TUNE_FIT_TILE_ROWS, 0) == 4096
    ev.close()


def _linear(p, n):
    t = 2 * np.pi * np.arange(n) / (n - 1)
    x = np.stack([np.sin((k + 1) * t) for k in range(p)], 1)
    y = x @ (1.0 + 0.31 * np.arange(p)) + (((np.arange(n) * 37) % 11) - 5.0) * 2e-4
    terms = [node(MUL, const(0.25 + 0.07 * k), var(k)) for k in range(p)]
    tree = terms[-1]
    for tt in reversed(terms[:-1]):
        tree = node(ADD, tt, tree)
    return x, y, tree


@pytest.mark.parametrize("tile,rows", [(2048, 4), (2048, 2), (8192, 4), (32768, 2), (65536, 4)])
def test_every_fit_shape_reaches_the_oracle_optimum(tile, rows):
    n = 20000
    for p in (1, 5, 12, 20):
        x, y, tree = _linear(p, n)
        ev = _evaluator(x, y)
        assert ev.set_tuning(TUNE_FIT_TILE_ROWS, tile) >= 0
        assert ev.set_tuning(TUNE_FIT_ROWS_PER_LANE, rows) >= 0
        outs, ms = ev.fit_batch([host.gpu_fit_plan(tree)], 8, 0)
        ev.close()
        assert ms >= 0 and outs[0].status == 0, (tile, rows, p, outs[0].status)
        want, _ = oracle.fit_weighted(tree, x, y, None, 8)
        got = oracle.sse(tree.with_constants(outs[0].c), x, y)
        _say(f"tile {tile} rows {rows} p {p}: rel {abs(got - want) / want:.2e}")
        assert abs(got - want) <= 1e-8 * want


def test_hetero_batch_is_consistent_across_shapes():
    wl = workloads.config4(nisland=3, pop=16, nmin=512, nmax=40000, iters=8)
    X, Y, sets = wl.x, wl.y, wl.rowsets
    plans, isl, met = [], [], []
    for t, i, m in zip(wl.trees, wl.island_of, wl.metrics):
        p = host.gpu_fit_plan(t)
        if p is None:
            continue
        plans.append(p)
        isl.append(int(i))
        met.append(int(m))
    ev = _evaluator(X[sets[0]], Y[sets[0]])
    ids = {0: 0}
    for s in range(1, len(sets)):
        xs, ys = X[sets[s]], Y[sets[s]]
        ids[s] = ev.add_rows(np.ascontiguousarray(xs.T, np.float32).reshape(-1), len(xs),
                             xs.shape[1], ys.astype(np.float32),
                             np.ascontiguousarray(xs.T).reshape(-1), ys, None)
        assert ids[s] > 0
    rowset = [ids[i] for i in isl]
    base, ms = ev.fit_batch_hetero(plans, rowset, met, 8, kmax=20)
    assert ms >= 0
    for tile, rows in ((2048, 4), (32768, 2)):
        ev.set_tuning(TUNE_FIT_TILE_ROWS, tile)
        ev.set_tuning(TUNE_FIT_ROWS_PER_LANE, rows)
        outs, ms = ev.fit_batch_hetero(plans, rowset, met, 8, kmax=20)
        assert ms >= 0
        rel = []
        for a, b in zip(outs, base):
            assert host.fit_status_usable(a.status) == host.fit_status_usable(b.status)
            if b.status in (0, 4) and a.status in (0, 4) and b.sse > 0:
                rel.append(abs(a.sse - b.sse) / b.sse)
        _say(f"tile {tile} rows {rows}: {len(rel)} compared, median rel {np.median(rel):.2e}")
        assert np.median(rel) <= 1e-6
    ev.close()


def test_eval_tiles_give_the_same_predictions_and_losses():
    wl = workloads.config2(n=20000, nprog=64)
    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    pop = host.Population()
    for t in wl.trees:
        pop.add(t)
    prog, off = pop.arrays()
    base_pred = ev.predict_words(prog, off) if hasattr(ev, "predict_words") else None
    base, ms = ev.eval_words(prog, off, 0)
    assert ms >= 0
    for tile in (1024, 8192, 16384):
        assert ev.set_tuning(TUNE_EVAL_TILE_ROWS, tile) >= 0
        out, ms = ev.eval_words(prog, off, 0)
        assert ms >= 0
        fin = np.isfinite(base) & np.isfinite(out)
        assert np.array_equal(np.isfinite(base), np.isfinite(out))
        assert np.allclose(out[fin], base[fin], rtol=1e-6, atol=0), tile
        if base_pred is not None:
            assert np.array_equal(ev.predict_words(prog, off), base_pred, equal_nan=True)
    ev.close()


def test_autotuner_times_applies_and_caches(tmp_path):
    wl = workloads.config3(n=16384, ncand=64, iters=8)
    ev = _evaluator(wl.x, wl.y)
    plans = [host.gpu_fit_plan(t) for t in wl.trees]
    cache = at.TuningCache(str(tmp_path / "tuning.json"))
    res = at.tune_fit(ev, plans, 8, 0, rows=[wl.n] * len(plans), sample=32, reps=1, cache=cache)
    assert res is not None and res.table
    _say(f"autotuned fit: {res.to_json()}")
    assert res.best in [s for s, _ in res.table]
    for k, v in res.best.items():
        assert ev.get_tuning(k) == v
    assert math.isfinite(res.best_ms) and res.best_ms <= min(ms for _, ms in res.table) * 1.02 + 1e-9
    outs, ms = ev.fit_batch(plans, 8, 0)
    assert ms >= 0 and all(o.status in (0, 4, 5, 6, 7, 8) for o in outs)
    ev.close()
