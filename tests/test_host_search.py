"""Host-side callers of the device fit on the CPU: the dispatch cost model
(dispatch.rs:766-903) and the search's admission / fit_pending / fallback
logic (search.rs:786-1051, 3019-3090), with a stub evaluator standing in for
the device so every status path runs without a GPU."""
import math

import numpy as np
import pytest

from rcsr_hip import dispatch as dsp
from rcsr_hip import search
from rcsr_hip.expr import ADD, EXP, MUL, SIN, const, node, var
from rcsr_hip.host import GpuFitOutput


def _cal(**kw):
    base = dict(r_eval=1e10, r_fit=100.0, bw=1e10, launch_s=1e-5, cpu_eval=1e8, cpu_fit=1.0,
                hysteresis=0.1)
    base.update(kw)
    return dsp.Calibration(**base)


def test_admission_count_matches_the_reference_rule():
    assert search.admission_count(400, 1000) == 100
    assert search.admission_count(64, 1000) == 32
    assert search.admission_count(10, 1000) == 10
    assert search.admission_count(400, 1 << 20) == 400
    assert search.admission_count(0, 10) == 1


def test_fit_route_follows_the_cost_model_and_hysteresis():
    d = dsp.Dispatcher(_cal())
    w = dsp.Workload.fit(1, 100_000, 4, 8)
    dec = d.plan(dsp.Stage.FIT, w)
    # device rate 100 fits/s x 24.6 (row scaling) >> cpu -> device
    assert dec.route == dsp.Route.DEVICE and dec.reason == dsp.Reason.COST_DEVICE
    c = dec.cost
    assert c.fit_s == pytest.approx(1.0 / (100.0 * dsp.FIT_REF_ROWS / 100_000))
    assert c.cpu_s == pytest.approx(1.0 / (1.0 * dsp.FIT_REF_ROWS / 100_000))
    # A calibration where the device is slower than the CPU by 5 %: within
    # the hold band after a device decision, outside it otherwise.
    slow = dsp.Dispatcher(_cal(r_fit=0.95, launch_s=0.0, bw=1e30))
    assert slow.plan(dsp.Stage.FIT, w).route == dsp.Route.CPU
    slow.last[dsp.Stage.FIT] = dsp.Route.DEVICE
    dec = slow.plan(dsp.Stage.FIT, w)
    assert dec.route == dsp.Route.DEVICE and dec.reason == dsp.Reason.HYSTERESIS_DEVICE
    # Row floor, missing calibration, refusal, CPU-contract stage.
    assert d.plan(dsp.Stage.FIT, dsp.Workload.fit(1, 50, 4, 8).with_min_rows(128)).reason == \
        dsp.Reason.TOO_FEW_ROWS
    assert dsp.Dispatcher(_cal(r_fit=0.0)).plan(dsp.Stage.FIT, w).reason == \
        dsp.Reason.NO_CALIBRATION
    w2 = dsp.Workload.fit(1, 100_000, 4, 8)
    w2.refusal = dsp.Reason.UNSUPPORTED_PLAN
    assert d.plan(dsp.Stage.FIT, w2).reason == dsp.Reason.UNSUPPORTED_PLAN
    assert d.plan(dsp.Stage.SELECT, w).reason == dsp.Reason.CPU_CONTRACT
    assert dsp.estimate_fit_bytes(3, 5) == 3 * (256 + 40)


def _islands():
    trees = [node(MUL, const(1.0), var(0)),                                  # k = 1
             node(ADD, node(MUL, const(2.0), var(0)), const(0.5)),           # k = 2
             node(SIN, var(0)),                                             # no constants
             node(MUL, const(0.3), node(EXP, node(MUL, const(0.2), var(0)))),  # k = 2
             node(ADD, node(MUL, const(1.5), var(0)), const(-1.0))]          # k = 2
    errs = [0.5, 0.1, 0.05, 0.3, 0.1]
    cx = [3, 5, 2, 6, 4]
    return [[search.Ind(t, e, c) for t, e, c in zip(trees, errs, cx)]]


def test_select_promoted_ranks_by_error_complexity_index_and_skips_constant_free():
    isl = _islands()
    pending = [(0, j) for j in range(5)]
    got = search.select_promoted(isl, pending, 1 << 20)
    assert got == [(0, 0), (0, 1), (0, 3), (0, 4)]  # tree 2 has nothing to fit
    # Narrow data: admission_count(5, rows) = 5, still sorted by index.
    assert search.select_promoted(isl, pending, 100) == [(0, 0), (0, 1), (0, 3), (0, 4)]


class _StubEv:
    def __init__(self, statuses, fail=False):
        self.statuses = statuses
        self.fail = fail
        self.calls = []

    def fit_batch(self, plans, iters, metric, kmax=None):
        self.calls.append((len(plans), iters, metric))
        if self.fail:
            return [], -1.0
        outs = []
        for p, st in zip(plans, self.statuses):
            c = p.c0 + 1.0
            if st == "nan":
                c = c * np.nan
                st = 0
            outs.append(GpuFitOutput(c, 1.0, st))
        return outs, 2.0


def test_fit_pending_applies_labels_and_falls_back():
    isl = _islands()
    admitted = [(0, 0), (0, 1), (0, 3), (0, 4)]
    ev = _StubEv([0, 4, 1, "nan"])
    refit = []

    def cpu_fit(t):
        refit.append(t)
        return t
    d = dsp.Dispatcher(_cal())
    rep = search.fit_pending(ev, isl, admitted, 100_000, 8, 0, d, cpu_fit)
    assert ev.calls == [(4, 8, 0)]
    assert rep.fit_gpu == 2 and rep.status_labels == {"rank_deficient": 1}
    causes = {(i, j): c for (i, j, c) in rep.fallbacks}
    assert causes == {(0, 3): search.FitFallbackCause.ILL_CONDITIONED,
                      (0, 4): search.FitFallbackCause.NON_FINITE}
    # Finite device constants are applied (status 1 included); NaN ones are not.
    assert isl[0][0].tree.constants() == [2.0]
    assert isl[0][3].tree.constants() == [1.3, 1.2]
    assert isl[0][4].tree.constants() == [1.5, -1.0]
    assert len(refit) == 2 and rep.fallback_bounded == 0
    led = d.ledger[dsp.Stage.FIT]
    assert led.calls == 4 and led.actual_s == pytest.approx(0.002)


def test_fit_pending_launch_failure_metric_and_bound():
    isl = _islands()
    admitted = [(0, 0), (0, 1), (0, 3)]
    rep = search.fit_pending(_StubEv([0, 0, 0], fail=True), isl, admitted, 100_000, 8, 0,
                             dsp.Dispatcher(_cal()), lambda t: t, fallback_max=2)
    assert [c for (_, _, c) in rep.fallbacks] == [search.FitFallbackCause.LAUNCH] * 3
    assert rep.fallback_refit == 2 and rep.fallback_bounded == 1
    rep = search.fit_pending(_StubEv([0, 0, 0]), _islands(), admitted, 100_000, 8, 5,
                             dsp.Dispatcher(_cal()), lambda t: t)
    assert {c for (_, _, c) in rep.fallbacks} == {search.FitFallbackCause.METRIC_NOT_NATIVE}
    # The cost model sends a slow-device shape to the CPU as a policy route.
    rep = search.fit_pending(_StubEv([0, 0, 0]), _islands(), admitted, 100_000, 8, 0,
                             dsp.Dispatcher(_cal(r_fit=1e-3)), lambda t: t)
    assert {c for (_, _, c) in rep.fallbacks} == {search.FitFallbackCause.DISPATCH_COST}
    # The row floor is 32 rows per constant: 40 rows admit k = 1, not k = 2.
    rep = search.fit_pending(_StubEv([0, 0, 0]), _islands(), admitted, 40, 8, 0,
                             dsp.Dispatcher(_cal(cpu_fit=1e-6)), lambda t: t)
    assert [c for (_, _, c) in rep.fallbacks] == [search.FitFallbackCause.TOO_FEW_ROWS] * 2
    assert rep.fit_gpu == 1
    # More than 20 constants: a plan rejection.
    big = node(MUL, const(1.0), var(0))
    for q in range(21):
        big = node(ADD, big, node(MUL, const(0.1 * q), var(0)))
    isl = [[search.Ind(big, 0.1, 50)]]
    rep = search.fit_pending(_StubEv([0]), isl, [(0, 0)], 100_000, 8, 0,
                             dsp.Dispatcher(_cal()), lambda t: t)
    assert rep.fallbacks == [(0, 0, search.FitFallbackCause.TOO_MANY_CONSTANTS)]


def test_device_rates_replace_only_usable_device_terms():
    # dispatch.rs:514-530: each finite positive probe value replaces the
    # calibration's device term; the CPU terms and hysteresis stay.
    cal = _cal()
    got = dsp.with_device_rates(cal, dsp.DeviceRates(r_eval=3e11, r_fit=50.0, bw=2e10,
                                                     launch_s=4e-6))
    assert (got.r_eval, got.r_fit, got.bw, got.launch_s) == (3e11, 50.0, 2e10, 4e-6)
    assert (got.cpu_eval, got.cpu_fit, got.hysteresis) == (cal.cpu_eval, cal.cpu_fit,
                                                           cal.hysteresis)
    bad = dsp.with_device_rates(cal, dsp.DeviceRates(r_eval=math.nan, r_fit=0.0, bw=-1.0,
                                                     launch_s=-1e-6))
    assert (bad.r_eval, bad.r_fit, bad.bw, bad.launch_s) == (cal.r_eval, cal.r_fit, cal.bw,
                                                             cal.launch_s)
    assert dsp.with_device_rates(cal, dsp.DeviceRates(1.0, 1.0, 1.0, 0.0)).launch_s == 0.0
