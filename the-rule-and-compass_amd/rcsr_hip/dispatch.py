"""The per-stage dispatch cost model that routes a fit or an evaluation to
the device or to the CPU (dispatch.rs:184-214, 330-395, 530-569, 766-903,
994-1002).

The reference keeps this model above the C ABI, unchanged by the device
arm; it is restated here so the Python host side (rcsr_hip.search) makes the
same route decisions the Rust search makes in front of librcsrhip.so:

    route = Device  iff  T_dev + T_fit + T_xfer + T_launch < T_cpu / (1 + h)

with a hold band T_cpu (1 + h) once the stage is already on the device, a
row floor (fewer rows than the normal equations need stay on the CPU) and a
ledger of predicted and measured seconds per stage.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from enum import Enum
from typing import Dict, Optional

# Reference shape of the calibrated fit rates (dispatch.rs:85-87).
FIT_REF_ROWS = 2_456_306.0
FIT_REF_CONSTANTS = 4.0
FIT_REF_SWEEPS = 8.0
# The reference's device anchors (dispatch.rs:69-81): row-evals/s, accepted
# fits/s at the reference shape, host-device bytes/s and launch seconds.
DEFAULT_R_EVAL = 20.0e9
DEFAULT_R_FIT = 1.0 / 0.0895
DEFAULT_BW = 12.0e9
DEFAULT_LAUNCH_S = 5.0e-6


class Stage(Enum):
    EVAL = "eval"
    FIT = "fit"
    SELECT = "select"  # a CPU-contract stage


class Route(Enum):
    DEVICE = "device"
    CPU = "cpu"
    SPLIT = "split"


class Reason(Enum):
    COST_DEVICE = "cost_device"
    COST_CPU = "cost_cpu"
    HYSTERESIS_DEVICE = "hysteresis_device"
    SPLIT_COST = "split_cost"
    CPU_CONTRACT = "cpu_contract"
    NO_DEVICE = "no_device"
    UNSUPPORTED_PLAN = "unsupported_plan"
    TOO_FEW_ROWS = "too_few_rows"
    TOO_MANY_CONSTANTS = "too_many_constants"
    LAUNCH = "launch"
    ILL_CONDITIONED = "ill_conditioned"
    NON_FINITE = "non_finite"
    METRIC_NOT_NATIVE = "metric_not_native"
    EXACTNESS = "exactness"
    NO_CALIBRATION = "no_calibration"


def _fdiv(a: float, b: float) -> float:
    """IEEE division as the reference's f64 arithmetic does it (a / 0 is
    +-inf, 0 / 0 is NaN) instead of Python's ZeroDivisionError."""
    if b == 0.0:
        return math.nan if a == 0.0 or math.isnan(a) else math.copysign(math.inf, a)
    return a / b


def estimate_eval_bytes(candidates: int) -> int:
    """dispatch.rs:994: program words and one result per candidate."""
    return max(candidates, 1) * (256 + 8)


def estimate_fit_bytes(candidates: int, constants: int) -> int:
    """dispatch.rs:999: constants and flags both ways per candidate."""
    return max(candidates, 1) * (256 + 8 * constants)


@dataclass
class Workload:
    candidates: int
    rows: int
    nodes: int = 1
    constants: int = 0
    sweeps: int = 0
    bytes: int = 0
    min_rows: int = 1
    cpu_s: Optional[float] = None
    device_s: Optional[float] = None
    split_fraction: Optional[float] = None
    refusal: Optional[Reason] = None

    @classmethod
    def eval(cls, candidates: int, rows: int, nodes: int) -> "Workload":
        return cls(candidates, rows, nodes=nodes, bytes=estimate_eval_bytes(candidates))

    @classmethod
    def fit(cls, candidates: int, rows: int, constants: int, sweeps: int) -> "Workload":
        return cls(candidates, rows, nodes=1, constants=constants, sweeps=sweeps,
                   bytes=estimate_fit_bytes(candidates, constants))

    def with_min_rows(self, rows: int) -> "Workload":
        self.min_rows = max(rows, 1)
        return self


@dataclass
class Calibration:
    """Measured rates: row-evals/s (device, CPU), fit candidates/s at the
    reference shape (device, CPU), host-device bytes/s, launch seconds and
    the hysteresis band."""
    r_eval: float
    r_fit: float
    bw: float
    launch_s: float
    cpu_eval: float
    cpu_fit: float
    hysteresis: float = 0.1

    def is_usable(self) -> bool:
        pos = [self.r_eval, self.r_fit, self.bw, self.cpu_eval, self.cpu_fit]
        return (all(math.isfinite(v) and v > 0.0 for v in pos) and math.isfinite(self.launch_s)
                and self.launch_s >= 0.0 and math.isfinite(self.hysteresis)
                and self.hysteresis >= 0.0)

    @staticmethod
    def _scale(rate: float, rows: int, constants: int, sweeps: int) -> float:
        return (rate * (FIT_REF_ROWS / max(rows, 1)) * (FIT_REF_CONSTANTS / max(constants, 1))
                * (FIT_REF_SWEEPS / max(sweeps, 1)))

    def fit_rate(self, rows: int, constants: int, sweeps: int) -> float:
        """Device fits per second at a (rows, constants, sweeps) shape."""
        return self._scale(self.r_fit, rows, constants, sweeps)

    def cpu_fit_rate(self, rows: int, constants: int, sweeps: int) -> float:
        """Exact CPU fits per second at the same shape."""
        return self._scale(self.cpu_fit, rows, constants, sweeps)


@dataclass
class DeviceRates:
    """dispatch.rs:642-660: device rates measured by a hardware probe:
    row-evals/s, fits/s at the reference shape (FIT_REF_ROWS rows,
    FIT_REF_CONSTANTS constants, FIT_REF_SWEEPS sweeps), host-device bytes/s
    and the launch overhead in seconds."""
    r_eval: float
    r_fit: float
    bw: float
    launch_s: float


def default_device_rates() -> DeviceRates:
    """dispatch.rs:649-658: the anchors, for a run without a probe."""
    return DeviceRates(DEFAULT_R_EVAL, DEFAULT_R_FIT, DEFAULT_BW, DEFAULT_LAUNCH_S)


def with_device_rates(cal: "Calibration", rates: DeviceRates) -> "Calibration":
    """dispatch.rs:514-530 Calibration::with_device_rates: each usable probe
    value replaces the calibration's device term; an unusable one (not
    finite, or not positive; a negative launch time) leaves it."""
    def usable(v, allow_zero=False):
        return math.isfinite(v) and (v >= 0.0 if allow_zero else v > 0.0)
    return Calibration(
        r_eval=rates.r_eval if usable(rates.r_eval) else cal.r_eval,
        r_fit=rates.r_fit if usable(rates.r_fit) else cal.r_fit,
        bw=rates.bw if usable(rates.bw) else cal.bw,
        launch_s=rates.launch_s if usable(rates.launch_s, True) else cal.launch_s,
        cpu_eval=cal.cpu_eval, cpu_fit=cal.cpu_fit, hysteresis=cal.hysteresis)


@dataclass
class Cost:
    row_evals: float
    device_s: float
    fit_s: float
    transfer_s: float
    launch_s: float
    cpu_s: float
    predicted_s: float = 0.0
    device_ratio: float = math.inf


@dataclass
class Decision:
    stage: Stage
    route: Route
    reason: Reason
    cost: Cost


@dataclass
class StageLedger:
    route: Route = Route.CPU
    reason: Reason = Reason.NO_CALIBRATION
    calls: int = 0
    candidates: int = 0
    predicted_s: float = 0.0
    actual_s: float = 0.0
    bytes: int = 0
    switches: int = 0


@dataclass
class Dispatcher:
    calibration: Calibration
    last: Dict[Stage, Route] = field(default_factory=dict)
    ledger: Dict[Stage, StageLedger] = field(default_factory=dict)

    def cost(self, stage: Stage, w: Workload) -> Cost:
        cal = self.calibration
        row_evals = float(max(w.candidates, 1) * max(w.rows, 1) * max(w.nodes, 1))
        if w.bytes == 0:
            transfer_s = 0.0
        elif cal.bw > 0.0:
            transfer_s = w.bytes / cal.bw
        else:
            transfer_s = math.inf
        launch_s = cal.launch_s
        if w.cpu_s is not None and w.device_s is not None:
            device_s, fit_s, cpu_s = w.device_s, 0.0, w.cpu_s
        elif stage == Stage.EVAL:
            device_s = row_evals / cal.r_eval if cal.r_eval > 0.0 else math.inf
            cpu_s = row_evals / cal.cpu_eval if cal.cpu_eval > 0.0 else math.inf
            fit_s = 0.0
        elif stage == Stage.FIT:
            cands = float(max(w.candidates, 1))
            device_s = 0.0
            fit_s = _fdiv(cands, cal.fit_rate(w.rows, w.constants, w.sweeps))
            cpu_s = _fdiv(cands, cal.cpu_fit_rate(w.rows, w.constants, w.sweeps))
        else:
            device_s = fit_s = cpu_s = 0.0
        total = device_s + fit_s + transfer_s + launch_s
        ratio = total / cpu_s if (math.isfinite(cpu_s) and cpu_s > 0.0) else math.inf
        return Cost(row_evals, device_s, fit_s, transfer_s, launch_s, cpu_s, 0.0, ratio)

    def _cost_route(self, stage: Stage, c: Cost):
        if not (math.isfinite(c.cpu_s) and c.cpu_s > 0.0):
            return Route.CPU, Reason.NO_CALIBRATION
        t_dev = c.device_s + c.fit_s + c.transfer_s + c.launch_s
        h = max(self.calibration.hysteresis, 0.0)
        base_device = t_dev < c.cpu_s * (1.0 / (1.0 + h))
        if self.last.get(stage) == Route.DEVICE and t_dev < c.cpu_s * (1.0 + h):
            return Route.DEVICE, (Reason.COST_DEVICE if base_device else Reason.HYSTERESIS_DEVICE)
        if base_device:
            return Route.DEVICE, Reason.COST_DEVICE
        return Route.CPU, Reason.COST_CPU

    def plan(self, stage: Stage, w: Workload) -> Decision:
        c = self.cost(stage, w)
        if w.refusal is not None:
            route, reason = Route.CPU, w.refusal
        elif w.rows < w.min_rows:
            route, reason = Route.CPU, Reason.TOO_FEW_ROWS
        elif stage == Stage.SELECT:
            route, reason = Route.CPU, Reason.CPU_CONTRACT
        elif not self.calibration.is_usable():
            route, reason = Route.CPU, Reason.NO_CALIBRATION
        elif w.split_fraction is not None and 0.0 < w.split_fraction < 1.0:
            route, reason = Route.SPLIT, Reason.SPLIT_COST
        else:
            route, reason = self._cost_route(stage, c)
        dev = c.device_s + c.fit_s + c.transfer_s + c.launch_s
        if route == Route.DEVICE:
            c.predicted_s = dev
        elif route == Route.CPU:
            c.predicted_s = c.cpu_s
        else:
            f = w.split_fraction if w.split_fraction is not None else 0.5
            c.predicted_s = f * dev + (1.0 - f) * c.cpu_s
        prev = self.last.get(stage)
        self.last[stage] = route
        e = self.ledger.setdefault(stage, StageLedger())
        if prev is not None and prev != route:
            e.switches += 1
        e.route, e.reason = route, reason
        return Decision(stage, route, reason, c)

    def record(self, d: Decision, actual_s: float, candidates: int, nbytes: int) -> None:
        e = self.ledger.setdefault(d.stage, StageLedger())
        e.route, e.reason = d.route, d.reason
        e.calls += 1
        e.candidates += candidates
        e.predicted_s += d.cost.predicted_s
        if math.isfinite(actual_s) and actual_s >= 0.0:
            e.actual_s += actual_s
        e.bytes += nbytes
