"""The launch-shape autotuner's search, margin rule, refusals, cache and
signatures on the CPU, with an evaluator stand-in whose device time is a
function of the shapes."""
import math

import numpy as np

from rcsr_hip import autotune as at
from rcsr_hip.host import TUNE_EVAL_TILE_ROWS, TUNE_FIT_ROWS_PER_LANE, TUNE_FIT_TILE_ROWS


class FakeEv:
    """get/set_tuning like the C ABI (0 = built-in; refuses what it should)."""

    def __init__(self, cost):
        self.t = {TUNE_FIT_TILE_ROWS: 0, TUNE_FIT_ROWS_PER_LANE: 0, TUNE_EVAL_TILE_ROWS: 0}
        self.cost = cost
        self.calls = 0

    def set_tuning(self, k, v):
        if k in (TUNE_FIT_TILE_ROWS, TUNE_EVAL_TILE_ROWS):
            if v != 0 and (v % 256 or v < 256 or v > 1 << 22):
                return -1
        elif v not in (0, 2, 4):
            return -1
        prev = self.t[k]
        self.t[k] = v
        return prev

    def get_tuning(self, k):
        return self.t[k]

    def run(self):
        self.calls += 1
        return self.cost(self.t)


def test_tune_picks_the_fastest_and_applies_it():
    def cost(t):
        tile = t[TUNE_FIT_TILE_ROWS] or 8192
        rows = t[TUNE_FIT_ROWS_PER_LANE] or 4
        return 100.0 + abs(math.log2(tile / 32768)) * 10 + (5 if rows == 2 else 0)
    ev = FakeEv(cost)
    res = at.tune(ev, ev.run, at.fit_space(4_000_000), reps=2)
    assert res.best == {TUNE_FIT_TILE_ROWS: 32768, TUNE_FIT_ROWS_PER_LANE: 4}
    assert ev.t[TUNE_FIT_TILE_ROWS] == 32768 and ev.t[TUNE_FIT_ROWS_PER_LANE] == 4
    assert res.best_ms == 100.0 and res.builtin_ms == 120.0
    assert len(res.table) == len(at.fit_space(4_000_000)) and ev.calls == 2 * len(res.table)
    assert res.to_json()["best"] == {"fit_tile_rows": 32768, "fit_rows_per_lane": 4}


def test_builtin_kept_within_the_margin_and_refusals_skipped():
    ev = FakeEv(lambda t: 100.0 if (t[TUNE_FIT_TILE_ROWS] or 8192) == 8192 else 99.0)
    res = at.tune(ev, ev.run, at.fit_space(1 << 20), margin=0.02)
    assert res.best[TUNE_FIT_TILE_ROWS] == 8192  # 1% faster is noise
    space = [{TUNE_FIT_TILE_ROWS: 300}, {TUNE_FIT_TILE_ROWS: 4096}]  # 300 is refused
    res = at.tune(ev, ev.run, space)
    assert [s for s, _ in res.table] == [{TUNE_FIT_TILE_ROWS: 4096}]


def test_failed_runs_restore_the_previous_shapes():
    ev = FakeEv(lambda t: -1.0)
    ev.set_tuning(TUNE_FIT_TILE_ROWS, 16384)
    assert at.tune(ev, ev.run, at.fit_space(1 << 20)) is None
    assert ev.t[TUNE_FIT_TILE_ROWS] == 16384
    # A setting whose runs partly fail keeps its successful timings.
    seq = iter([-1.0, 7.0])
    ev = FakeEv(lambda t: next(seq))
    res = at.tune(ev, ev.run, [{TUNE_EVAL_TILE_ROWS: 2048}], reps=2)
    assert res.best_ms == 7.0


def test_spaces_cover_the_largest_candidate_once():
    sp = at.fit_space(5000)
    tiles = sorted({s[TUNE_FIT_TILE_ROWS] for s in sp})
    assert tiles == [2048, 4096, 8192]  # 8192 covers 5000 rows and is built in
    assert {s[TUNE_FIT_ROWS_PER_LANE] for s in sp} == {2, 4}
    assert [s[TUNE_EVAL_TILE_ROWS] for s in at.eval_space(1 << 20)] == list(at.EVAL_TILES)
    assert [s[TUNE_EVAL_TILE_ROWS] for s in at.eval_space(1500)] == [1024, 2048, 4096]


def test_signature_buckets_and_cache_round_trip(tmp_path):
    a = at.signature("fit", [262144] * 1024, [10] * 1024, [0] * 1024)
    b = at.signature("fit", [300000] * 1024, [11] * 1024, [0] * 1024)
    c = at.signature("fit", [262144] * 1024, [10] * 1024, [4] * 1024)
    assert a == b and a != c
    assert at.signature("fit", [512, 4_000_000], [3, 4], [0, 4]).endswith(":mixed")
    path = tmp_path / "sub" / "tuning.json"
    cache = at.TuningCache(str(path))
    cache.put(a, {TUNE_FIT_TILE_ROWS: 16384, TUNE_FIT_ROWS_PER_LANE: 4})
    again = at.TuningCache(str(path))
    assert again.get(a) == {TUNE_FIT_TILE_ROWS: 16384, TUNE_FIT_ROWS_PER_LANE: 4}
    assert again.get(c) is None


class Plan:
    def __init__(self, k):
        self.nconst = k


def test_tune_fit_uses_the_cache_and_an_even_sample(tmp_path):
    ev = FakeEv(lambda t: 10.0)
    runs = []

    def fit_batch(plans, iters, metric):
        runs.append(len(plans))
        return [], ev.run()
    ev.fit_batch = fit_batch
    plans = [Plan(k % 20 + 1) for k in range(1000)]
    cache = at.TuningCache(str(tmp_path / "t.json"))
    res = at.tune_fit(ev, plans, 30, 0, rows=[262144] * 1000, sample=100, cache=cache, reps=1)
    assert res is not None and set(runs) == {100}
    n = len(runs)
    ev.set_tuning(TUNE_FIT_TILE_ROWS, 0)
    res2 = at.tune_fit(ev, plans, 30, 0, rows=[262144] * 1000, sample=100, cache=cache)
    assert len(runs) == n and math.isnan(res2.best_ms)  # applied from the cache, no timing
    assert ev.t[TUNE_FIT_TILE_ROWS] == res.best[TUNE_FIT_TILE_ROWS]


def test_tune_fit_confirms_a_sample_winner_on_the_whole_batch():
    # On the 100-candidate sample 4096-row items win; on the whole batch
    # they lose to the built-in shape, which confirmation then keeps.  When
    # the whole batch agrees with the sample, the winner stays.
    for full_prefers_builtin in (True, False):
        ev = FakeEv(None)
        size = {"n": 0}

        def cost(t, size=size, flip=full_prefers_builtin):
            tile = t[TUNE_FIT_TILE_ROWS] or 8192
            rows = t[TUNE_FIT_ROWS_PER_LANE] or 4
            base = 10.0 if size["n"] < 1000 else 40.0
            if rows == 2:
                return base * 1.5
            if tile == 4096:
                return base * (0.9 if (size["n"] < 1000 or not flip) else 1.1)
            return base
        ev.cost = cost

        def fit_batch(plans, iters, metric):
            size["n"] = len(plans)
            return [], ev.run()
        ev.fit_batch = fit_batch
        plans = [Plan(k % 20 + 1) for k in range(1000)]
        res = at.tune_fit(ev, plans, 30, 0, rows=[262144] * 1000, sample=100, reps=1,
                          confirm=True, space=[{TUNE_FIT_TILE_ROWS: t, TUNE_FIT_ROWS_PER_LANE: r}
                                               for t in (4096, 8192) for r in (4, 2)])
        assert res is not None
        want = 8192 if full_prefers_builtin else 4096
        assert res.best[TUNE_FIT_TILE_ROWS] == want
        assert ev.t[TUNE_FIT_TILE_ROWS] == want
        assert res.builtin_ms == 40.0
        # The whole-batch timings of both shapes are reported.
        conf = {s[TUNE_FIT_TILE_ROWS]: ms for s, ms in res.confirm_table}
        assert conf == {8192: 40.0, 4096: 44.0 if full_prefers_builtin else 36.0}
        assert len(res.to_json()["confirm_table"]) == 2
