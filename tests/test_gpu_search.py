"""search.fit_pending (search.rs:786-1051) driving the real library: the
admitted offspring of C1 islands are planned, routed by the cost model, fitted
in one rc_fit_batch on the evaluator's resident rows, labelled and refit on
the CPU where the device result is unusable.  The applied constants are the
device's own, bit for bit, and every fallback is the oracle's exact refit."""
import sys

import numpy as np
import pytest

import oracle
from rcsr_hip import dispatch as dsp
from rcsr_hip import host, search, workloads

pytestmark = pytest.mark.gpu


def test_fit_pending_on_the_device_applies_its_constants_and_refits_the_rest():
    wl = workloads.config1()
    x, y = np.asarray(wl.x), np.asarray(wl.y)
    n = len(y)
    rng = np.random.default_rng(3)
    islands = [[search.Ind(t, float(rng.uniform(0.1, 2.0)), float(len(t.constants()) + 1))
                for t in wl.trees[200 * i:200 * i + 200]] for i in range(4)]
    pending = [(i, j) for i in range(4) for j in range(200)]
    admitted = search.select_promoted(islands, pending, n)
    print(f"fit_pending: {len(admitted)} admitted of {len(pending)}", flush=True)
    assert 0 < len(admitted) <= len(pending)
    before = {(i, j): islands[i][j].tree for (i, j) in admitted}
    ev = host.GpuEvaluator(x.T.astype(np.float32).reshape(-1), n, 1, y.astype(np.float32))
    assert ev.set_fit_data(x.T.reshape(-1), y)
    # Device rates far above the CPU's: every supported plan goes to the device.
    cal = dsp.Calibration(r_eval=1e11, r_fit=1e4, bw=1e10, launch_s=1e-5, cpu_eval=1e8,
                          cpu_fit=1.0)
    d = dsp.Dispatcher(cal)
    refits = []

    def cpu_fit(t):
        loss, c = oracle.fit_weighted(t, x, y, None, 8)
        refits.append(t)
        return t.with_constants(list(c)) if len(c) else t
    rep = search.fit_pending(ev, islands, admitted, n, 8, 0, d, cpu_fit)
    # The same candidates fitted directly through the ABI.
    planned = [(i, j) for (i, j) in admitted if host.gpu_fit_plan(before[(i, j)]) is not None]
    outs, ms = ev.fit_batch([host.gpu_fit_plan(before[k]) for k in planned], 8, 0, kmax=20)
    ev.close()
    assert ms >= 0
    fell = {(i, j): c for (i, j, c) in rep.fallbacks}
    print(f"fit_pending: {rep.fit_gpu} fitted on the device, {len(fell)} fallbacks "
          f"{sorted(set(c.value for c in fell.values()))}, labels {dict(rep.status_labels)}",
          flush=True)
    assert rep.fit_gpu + len(fell) == len(admitted)
    assert rep.fit_gpu > 0.5 * len(admitted)
    for k, o in zip(planned, outs):
        got = islands[k[0]][k[1]].tree
        usable = host.fit_status_usable(o.status) and np.all(np.isfinite(o.c))
        if usable:
            assert k not in fell
            assert np.array_equal(np.asarray(got.constants()), o.c)
        else:
            assert k in fell
    assert len(refits) == len(fell)
    assert d.ledger[dsp.Stage.FIT].calls == len(admitted) - sum(
        1 for c in fell.values() if c in (search.FitFallbackCause.UNSUPPORTED_PLAN,
                                          search.FitFallbackCause.TOO_MANY_CONSTANTS))


def test_device_rate_probe_feeds_the_cost_model():
    # dispatch.rs:514-530, 642-660: the probe's four rates are finite and
    # positive, replace the calibration's device terms, and at the
    # reference's calibration shape the device's fit rate is the one the
    # model then uses.
    rates = host.probe_device_rates(rows=262144, fits=16, programs=64)
    print(f"probe: {rates}", flush=True)
    for v in (rates.r_eval, rates.r_fit, rates.bw):
        assert np.isfinite(v) and v > 0
    assert np.isfinite(rates.launch_s) and rates.launch_s >= 0
    anchors = dsp.default_device_rates()
    cal = dsp.with_device_rates(dsp.Calibration(
        r_eval=anchors.r_eval, r_fit=anchors.r_fit, bw=anchors.bw, launch_s=anchors.launch_s,
        cpu_eval=1.5e8, cpu_fit=1.0 / 6.232), rates)
    assert (cal.r_eval, cal.r_fit, cal.bw, cal.launch_s) == (rates.r_eval, rates.r_fit, rates.bw,
                                                             rates.launch_s)
    assert cal.fit_rate(int(dsp.FIT_REF_ROWS), 4, 8) == pytest.approx(rates.r_fit, rel=1e-12)
