"""The search-side callers of the batched device fit: admission of pending
offspring, the per-candidate dispatch plan, applying device results, status
labels and causes, and the bounded exact CPU fallback (search.rs:786-1051,
1562-1568, 2840-2959, 3006-3090).

They sit above the C ABI and are unchanged by the device arm; this module
restates them for the Python host side so a caller that drives
librcsrhip.so from Python routes, applies and falls back exactly as the
Rust search does.  The exact CPU fitter is the caller's (the reference's
fit_weighted_metric); it is passed in, never provided here.
"""
from __future__ import annotations

import time
from collections import Counter
from dataclasses import dataclass, field
from enum import Enum
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import dispatch as dsp
from .host import GpuFitReject, fit_status_usable, gpu_fit_plan_reason

WIDE_ADMISSION_ROWS = 1 << 20

# Metrics with a native device fit objective in rc_fit_batch (RC_LOSS_*).
NATIVE_FIT_METRICS = (0, 1, 4)


class FitFallbackCause(Enum):
    UNSUPPORTED_PLAN = "unsupported_plan"
    TOO_FEW_ROWS = "too_few_rows"
    TOO_MANY_CONSTANTS = "too_many_constants"
    METRIC_NOT_NATIVE = "metric_not_native"
    ILL_CONDITIONED = "ill_conditioned"
    NON_FINITE = "non_finite"
    LAUNCH = "launch"
    DISPATCH_COST = "dispatch_cost"


@dataclass
class Ind:
    """One individual of an island: its tree, screened error and complexity."""
    tree: object
    error: float = float("nan")
    complexity: float = 0.0


def admission_count(island_pop: int, train_rows: int) -> int:
    """search.rs:3019: every candidate on wide data, else the best quarter
    (at least 32) of the island."""
    if train_rows >= WIDE_ADMISSION_ROWS:
        return max(island_pop, 1)
    return max(min(max(island_pop // 4, 32), island_pop), 1)


def select_promoted(islands: Sequence[Sequence[Ind]], pending: Sequence[Tuple[int, int]],
                    train_rows: int) -> List[Tuple[int, int]]:
    """search.rs:3038-3090: per island, the best admission_count pending
    candidates by (screened error, complexity, index); candidates without
    constants are never admitted."""
    promoted = []
    for i, pop in enumerate(islands):
        ranked = []
        for (isl, j) in pending:
            if isl != i:
                continue
            if gpu_fit_plan_reason(pop[j].tree) == GpuFitReject.NO_CONSTANTS:
                continue
            ranked.append((j, pop[j].error, pop[j].complexity))

        def key(t):
            err = t[1] if not np.isnan(t[1]) else float("inf")
            cx = t[2] if not np.isnan(t[2]) else float("inf")
            return (err, cx, t[0])
        ranked.sort(key=key)
        take = min(admission_count(len(pop), train_rows), len(ranked))
        promoted.extend((i, j) for (j, _, _) in ranked[:take])
    promoted.sort()
    return promoted


def plan_fallback_cause(reason: GpuFitReject) -> FitFallbackCause:
    """search.rs:2926."""
    if reason == GpuFitReject.TOO_MANY_CONSTANTS:
        return FitFallbackCause.TOO_MANY_CONSTANTS
    return FitFallbackCause.UNSUPPORTED_PLAN


def status_fallback_cause(status: int) -> FitFallbackCause:
    """search.rs:2952: the kernel's failing statuses."""
    if status == 1:
        return FitFallbackCause.ILL_CONDITIONED
    if status == 2:
        return FitFallbackCause.NON_FINITE
    return FitFallbackCause.UNSUPPORTED_PLAN


def gpu_status_label(status: int) -> Optional[str]:
    """search.rs:2939: the ledger label of a finite device-handled fit."""
    return {4: "rank_deficient", 5: "recovered", 6: "no_improvement", 7: "start_unusable",
            8: "recovered_skipped"}.get(status)


def fallback_batch_split(total: int, cap: int) -> Tuple[int, int]:
    """search.rs:1562: (refit on the CPU, left bounded)."""
    if cap == 0:
        return total, 0
    return min(total, cap), max(total - cap, 0)


@dataclass
class FitReport:
    attempts: int = 0
    fit_gpu: int = 0
    fallbacks: List[Tuple[int, int, FitFallbackCause]] = field(default_factory=list)
    status_labels: Counter = field(default_factory=Counter)
    fallback_bounded: int = 0
    fallback_refit: int = 0
    device_ms: float = 0.0
    fallback_s: float = 0.0


def fit_pending(ev, islands: List[List[Ind]], admitted: Sequence[Tuple[int, int]],
                train_rows: int, iters: int, metric: int, dispatcher: dsp.Dispatcher,
                cpu_fit: Callable[[object], object], fallback_max: int = 0,
                kmax: int = 20) -> FitReport:
    """search.rs:786-1051 for the LM arms: plan each admitted candidate,
    route it through the cost model, run the device batch (ev.fit_batch on
    the evaluator's resident rows), apply finite constants, label statuses,
    and refit the fallbacks (at most fallback_max when it is not 0) with the
    caller's exact CPU fitter (cpu_fit(tree) -> tree)."""
    rep = FitReport(attempts=len(admitted))
    jobs, decisions, fallbacks = [], [], []
    native = metric in NATIVE_FIT_METRICS
    for (i, j) in admitted:
        reason = gpu_fit_plan_reason(islands[i][j].tree)
        if not native:
            fallbacks.append((i, j, FitFallbackCause.METRIC_NOT_NATIVE))
            continue
        if isinstance(reason, GpuFitReject):
            fallbacks.append((i, j, plan_fallback_cause(reason)))
            continue
        plan = reason
        w = dsp.Workload.fit(1, train_rows, plan.nconst, iters).with_min_rows(32 * plan.nconst)
        d = dispatcher.plan(dsp.Stage.FIT, w)
        if d.route == dsp.Route.DEVICE:
            jobs.append((i, j, plan))
            decisions.append(d)
        else:
            cause = {dsp.Reason.TOO_FEW_ROWS: FitFallbackCause.TOO_FEW_ROWS,
                     dsp.Reason.TOO_MANY_CONSTANTS: FitFallbackCause.TOO_MANY_CONSTANTS,
                     dsp.Reason.UNSUPPORTED_PLAN: FitFallbackCause.UNSUPPORTED_PLAN}.get(
                         d.reason, FitFallbackCause.DISPATCH_COST)
            fallbacks.append((i, j, cause))
            dispatcher.record(d, 0.0, 1, w.bytes)
    if jobs:
        outs, ms = ev.fit_batch([p for (_, _, p) in jobs], iters, metric, kmax=kmax)
        if ms >= 0 and len(outs) == len(jobs):
            rep.device_ms += ms
            for (i, j, plan), res in zip(jobs, outs):
                finite = bool(np.all(np.isfinite(res.c)))
                usable = fit_status_usable(res.status)
                if finite:
                    islands[i][j].tree = islands[i][j].tree.with_constants(list(res.c))
                if usable and finite:
                    rep.fit_gpu += 1
                    label = gpu_status_label(res.status)
                    if label:
                        rep.status_labels[label] += 1
                else:
                    cause = (FitFallbackCause.NON_FINITE if not finite
                             else status_fallback_cause(res.status))
                    fallbacks.append((i, j, cause))
        else:
            fallbacks.extend((i, j, FitFallbackCause.LAUNCH) for (i, j, _) in jobs)
        per = max(ms, 0.0) / 1000.0 / len(jobs)
        for d, (_, _, plan) in zip(decisions, jobs):
            dispatcher.record(d, per, 1, dsp.estimate_fit_bytes(1, plan.nconst))
    t0 = time.perf_counter()
    refit, bounded = fallback_batch_split(len(fallbacks), fallback_max)
    rep.fallback_bounded = bounded
    for (i, j, _) in fallbacks[:refit]:
        islands[i][j].tree = cpu_fit(islands[i][j].tree)
    rep.fallback_refit = refit
    rep.fallback_s = time.perf_counter() - t0
    rep.fallbacks = fallbacks
    return rep
