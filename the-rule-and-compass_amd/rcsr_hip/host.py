"""Host-side mirror of the reference's device-arm interface.

Names, argument meaning and error behaviour follow the reference's Rust host
layer so callers and tests read alike:

  compile_program / Population      gpu.rs:198-218, 972-1068
  gpu_fit_plan / GpuFitReject       fit.rs:958-1056
  GpuEvaluator                      gpu.rs:1182-1969 (new, set_fit_data,
                                    eval_batch, score_batch, predict,
                                    fit_batch, GpuFitOutput gpu.rs:828)
  fit_status_usable                 fit/gpu.rs:103
  BatchedGpu / fit_batch_verified   fit/gpu.rs:113-317, 456-654

The program builders run in C++ inside librcsrhip.so (rcsr_host.h); this
module only marshals arrays.  The exact CPU arm that the verifier falls back
to is the caller's (the reference keeps it unchanged); it is passed in as an
object with `sse`, `metric_score` and `fit_metric` methods.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import lib as _L

HOST_MAX_STACK = 56          # gpu.rs:134
GPU_FIT_MAX_CONSTANTS = 20   # fit.rs:851
GPU_FIT_MAX_NODES = 96       # fit.rs:854

METRIC_SQUARED, METRIC_LOG_SQUASHED, METRIC_ABSOLUTE = 0, 1, 4

FIT_MODE_MIXED, FIT_MODE_FP64 = 0, 1
# rc_set_tuning keys (include/rcsr_hip.h).
TUNE_FIT_TILE_ROWS, TUNE_FIT_ROWS_PER_LANE, TUNE_EVAL_TILE_ROWS = 1, 2, 3


class CompileError(Exception):
    """gpu.rs:142 CompileError: Unsupported(op) or StackOverflow{peak, limit}."""

    def __init__(self, kind: str, peak: int = 0):
        super().__init__(kind if kind != "stack_overflow" else
                         f"postfix program needs {peak} stack slots, the kernel allows "
                         f"{HOST_MAX_STACK}")
        self.kind = kind
        self.peak = peak


def _tree_arrays(tree) -> Tuple[np.ndarray, np.ndarray]:
    ops, vals = tree.preorder()
    return np.ascontiguousarray(ops, np.int32), np.ascontiguousarray(vals, np.float64)


def compile_program(tree) -> np.ndarray:
    """Postfix words, HALT-terminated (gpu.rs:198)."""
    L = _L.load()
    ops, vals = _tree_arrays(tree)
    cap = 16 * len(ops) + 64
    out = np.zeros(cap, np.uint32)
    peak = C.c_int(0)
    n = L.rch_compile_program(_L.ptr(ops, _L.i32p), _L.ptr(vals, _L.f64p), len(ops),
                              _L.ptr(out, _L.u32p), cap, C.byref(peak))
    if n == -1:
        raise CompileError("unsupported")
    if n == -2:
        raise CompileError("stack_overflow", peak.value)
    if n < 0:
        raise CompileError("malformed")
    return out[:n].copy()


class Population:
    """Packed postfix programs plus offsets (gpu.rs:972-1068)."""

    def __init__(self):
        self.prog: List[np.ndarray] = []
        self.offsets = [0]
        self.unsupported = 0

    def add(self, tree) -> bool:
        try:
            w = compile_program(tree)
        except CompileError:
            self.unsupported += 1
            return False
        self.prog.append(w)
        self.offsets.append(self.offsets[-1] + len(w))
        return True

    def __len__(self) -> int:
        return len(self.prog)

    def arrays(self) -> Tuple[np.ndarray, np.ndarray]:
        prog = np.concatenate(self.prog).astype(np.uint32) if self.prog else np.zeros(1, np.uint32)
        return np.ascontiguousarray(prog), np.asarray(self.offsets, np.uint32)


def _i128(hi, lo) -> int:
    """Two's-complement 128-bit value of (hi, lo) u64 halves."""
    u = (int(hi) << 64) | int(lo)
    return u - (1 << 128) if u >> 127 else u


class GpuFitReject(Enum):
    """fit.rs:958-983."""
    NO_CONSTANTS = "no_constants"
    TOO_MANY_CONSTANTS = "too_many_constants"
    TOO_MANY_NODES = "too_many_nodes"
    SHARED_CONSTANTS = "shared_constants"
    UNSUPPORTED_OP = "unsupported_op"


_REJECT = {-1: GpuFitReject.NO_CONSTANTS, -2: GpuFitReject.TOO_MANY_CONSTANTS,
           -3: GpuFitReject.TOO_MANY_NODES, -5: GpuFitReject.UNSUPPORTED_OP}


@dataclass
class GpuFitPlan:
    """fit.rs:929-950: 4-word post-order records and the start constants."""
    words: np.ndarray
    nconst: int
    c0: np.ndarray
    nodes: int


def gpu_fit_plan_reason(tree):
    """Returns a GpuFitPlan or a GpuFitReject (fit.rs:985)."""
    L = _L.load()
    ops, vals = _tree_arrays(tree)
    cap = 4 * (len(ops) + 1)
    words = np.zeros(cap, np.uint32)
    c0 = np.zeros(GPU_FIT_MAX_CONSTANTS + 1)
    nodes = C.c_int(0)
    k = L.rch_fit_plan(_L.ptr(ops, _L.i32p), _L.ptr(vals, _L.f64p), len(ops),
                       _L.ptr(words, _L.u32p), cap, _L.ptr(c0, _L.f64p), len(c0), C.byref(nodes))
    if k < 0:
        if k in _REJECT:
            return _REJECT[k]
        raise ValueError("malformed tree")
    return GpuFitPlan(words[:4 * nodes.value].copy(), k, c0[:k].copy(), nodes.value)


def gpu_fit_plan(tree) -> Optional[GpuFitPlan]:
    r = gpu_fit_plan_reason(tree)
    return r if isinstance(r, GpuFitPlan) else None


def fit_status_usable(status: int) -> bool:
    """fit/gpu.rs:103: statuses 0 and 4-8 carry a finite answer."""
    return status in (0, 4, 5, 6, 7, 8)


@dataclass(slots=True)
class GpuFitOutput:
    """gpu.rs:828."""
    c: np.ndarray
    sse: float
    status: int


def _fit_outputs(plans, outc, outsse, outst) -> List[GpuFitOutput]:
    """One output per plan: its constants (a view of its row of outc), loss
    and status."""
    rows = list(outc)
    sse = outsse.tolist()
    st = outst.tolist()
    return [GpuFitOutput(rows[i][:p.nconst], sse[i], st[i]) for i, p in enumerate(plans)]


def pack_plans(plans: Sequence[GpuFitPlan], kmax: Optional[int] = None):
    """The batch arrays of rc_fit_batch: concatenated records, their word
    offsets, constant counts and the start constants padded to kmax."""
    n = len(plans)
    pcount = np.fromiter((p.nconst for p in plans), np.int32, count=n)
    kmax = kmax or max(1, int(pcount.max()))
    ws = [p.words for p in plans]
    words = np.concatenate(ws).astype(np.uint32, copy=False)
    off = np.zeros(n + 1, np.uint32)
    off[1:] = np.cumsum(np.fromiter((len(w) for w in ws), np.int64, count=n))
    c0 = np.zeros((n, kmax))
    if int(pcount.sum()) > 0:
        c0[np.arange(kmax)[None, :] < pcount[:, None]] = np.concatenate([p.c0 for p in plans])
    return np.ascontiguousarray(words), off, pcount, c0, kmax


class GpuEvaluator:
    """Owns one device context with a resident dataset (gpu.rs:1182).

    `cols` is column-major f32 (cols[var * n + row]), like the reference.
    """

    def __init__(self, cols, n: int, v: int, y):
        self._L = _L.load()
        self.n, self.v = int(n), int(v)
        self._cols = np.ascontiguousarray(cols, np.float32)
        self._y = np.ascontiguousarray(y, np.float32)
        self.ctx = self._L.rc_create(_L.ptr(self._cols, _L.f32p), self.n, self.v,
                                     _L.ptr(self._y, _L.f32p))
        if not self.ctx:
            raise RuntimeError("rc_create failed (no usable device?)")

    @classmethod
    def new(cls, cols, n, v, y) -> Optional["GpuEvaluator"]:
        try:
            return cls(cols, n, v, y)
        except RuntimeError:
            return None

    def close(self):
        if self.ctx:
            self._L.rc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- data -------------------------------------------------------------
    def set_fit_data(self, cols64, y64, w=None) -> bool:
        c64 = np.ascontiguousarray(cols64, np.float64)
        y64 = np.ascontiguousarray(y64, np.float64)
        ww = None if w is None or len(w) == 0 else np.ascontiguousarray(w, np.float64)
        return self._L.rc_set_fit_data(self.ctx, _L.ptr(c64, _L.f64p), _L.ptr(y64, _L.f64p),
                                       _L.ptr(ww, _L.f64p)) == 0

    def add_rows(self, cols, n, v, y, cols64=None, y64=None, w=None) -> int:
        cols = np.ascontiguousarray(cols, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        c64 = None if cols64 is None else np.ascontiguousarray(cols64, np.float64)
        yy = None if y64 is None else np.ascontiguousarray(y64, np.float64)
        ww = None if w is None else np.ascontiguousarray(w, np.float64)
        return self._L.rc_add_rows(self.ctx, _L.ptr(cols, _L.f32p), int(n), int(v),
                                   _L.ptr(y, _L.f32p), _L.ptr(c64, _L.f64p), _L.ptr(yy, _L.f64p),
                                   _L.ptr(ww, _L.f64p))

    def set_score_weights(self, w=None) -> bool:
        """Per-row weights of the classification scores (None clears them)."""
        ww = None if w is None else np.ascontiguousarray(w, np.float64)
        return self._L.rc_set_score_weights(self.ctx, _L.ptr(ww, _L.f64p)) == 0

    def set_fit_sign(self, sign=None) -> bool:
        """Label signs t = 2l - 1 of the classification scores (None clears)."""
        ss = None if sign is None else np.ascontiguousarray(sign, np.float64)
        return self._L.rc_set_fit_sign(self.ctx, _L.ptr(ss, _L.f64p)) == 0

    def set_tuning(self, key: int, value: int) -> int:
        """Set a launch shape (TUNE_*; 0 = the built-in choice); returns the
        previous value, -1 when the key or value is refused."""
        return int(self._L.rc_set_tuning(self.ctx, int(key), int(value)))

    def get_tuning(self, key: int) -> int:
        return int(self._L.rc_get_tuning(self.ctx, int(key)))

    def set_fit_mode(self, mode: int) -> int:
        return self._L.rc_set_fit_mode(self.ctx, int(mode))

    # -- evaluation -------------------------------------------------------
    def eval_batch(self, trees, loss: int = 0):
        """Per-tree f32 loss sums (None for a tree the kernel cannot run)."""
        pop = Population()
        ok = [pop.add(t) for t in trees]
        if len(pop) == 0:
            return [None] * len(trees), 0.0
        prog, off = pop.arrays()
        out = np.zeros(len(pop), np.float32)
        self._L.rc_stats_begin(self.ctx)
        if self._L.rc_set_programs(self.ctx, _L.ptr(prog, _L.u32p), _L.ptr(off, _L.u32p),
                                   len(pop)) != 0:
            return [None] * len(trees), -1.0
        ms = self._L.rc_eval(self.ctx, _L.ptr(out, _L.f32p), int(loss))
        it = iter(out.tolist())
        return [next(it) if k else None for k in ok], ms

    def eval_words(self, prog: np.ndarray, off: np.ndarray, loss: int = 0):
        prog = np.ascontiguousarray(prog, np.uint32)
        off = np.ascontiguousarray(off, np.uint32)
        nprog = len(off) - 1
        out = np.zeros(nprog, np.float32)
        if self._L.rc_set_programs(self.ctx, _L.ptr(prog, _L.u32p), _L.ptr(off, _L.u32p),
                                   nprog) != 0:
            return None, -1.0
        ms = self._L.rc_eval(self.ctx, _L.ptr(out, _L.f32p), int(loss))
        return out, ms

    def score_words(self, prog, off, kind: int, pcount=None):
        prog = np.ascontiguousarray(prog, np.uint32)
        off = np.ascontiguousarray(off, np.uint32)
        nprog = len(off) - 1
        out = np.zeros(nprog)
        st = np.zeros(nprog, np.int32)
        pc = None if pcount is None else np.ascontiguousarray(pcount, np.int32)
        ms = self._L.rc_score_batch(self.ctx, _L.ptr(prog, _L.u32p), _L.ptr(off, _L.u32p), nprog,
                                    int(kind), _L.ptr(pc, _L.i32p), _L.ptr(out, _L.f64p),
                                    _L.ptr(st, _L.i32p))
        return out, st, ms

    def score_batch(self, trees, kind: int):
        pop = Population()
        ok = [pop.add(t) for t in trees]
        prog, off = pop.arrays()
        out, st, ms = self.score_words(prog, off, kind)
        it = iter(zip(out.tolist(), st.tolist()))
        return [next(it) if k else None for k in ok], ms

    def predict_words(self, prog, off) -> Optional[np.ndarray]:
        prog = np.ascontiguousarray(prog, np.uint32)
        off = np.ascontiguousarray(off, np.uint32)
        nprog = len(off) - 1
        out = np.zeros((nprog, self.n), np.float32)
        rc = self._L.rc_predict(self.ctx, _L.ptr(prog, _L.u32p), _L.ptr(off, _L.u32p), nprog,
                                _L.ptr(out, _L.f32p))
        return None if rc != 0 else out

    def predict(self, trees) -> Optional[np.ndarray]:
        pop = Population()
        for t in trees:
            if not pop.add(t):
                raise CompileError("unsupported")
        prog, off = pop.arrays()
        return self.predict_words(prog, off)

    # -- fit --------------------------------------------------------------
    def fit_batch(self, plans: Sequence[GpuFitPlan], iters: int, metric: int = 0,
                  kmax: Optional[int] = None, c0=None, gauge=None):
        """gpu.rs:1909: returns ([GpuFitOutput], ms); ms < 0 discards all.
        gauge: per plan the (scale, offset) constant slots the correlation
        metrics (2, 3) freeze, -1 for none."""
        words, off, pcount, c0p, kmax = pack_plans(plans, kmax)
        if c0 is not None:
            c0p = np.ascontiguousarray(c0, np.float64)
        n = len(plans)
        outc = np.zeros((n, kmax))
        outsse = np.zeros(n)
        outst = np.zeros(n, np.int32)
        gauge = (np.full(2 * n, -1, np.int32) if gauge is None
                 else np.ascontiguousarray(np.asarray(gauge, np.int32).reshape(-1)))
        ms = self._L.rc_fit_batch(self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p),
                                  _L.ptr(pcount, _L.i32p), n, kmax, _L.ptr(c0p, _L.f64p),
                                  _L.ptr(gauge, _L.i32p), int(iters), int(metric),
                                  _L.ptr(outc, _L.f64p), _L.ptr(outsse, _L.f64p),
                                  _L.ptr(outst, _L.i32p))
        return _fit_outputs(plans, outc, outsse, outst), ms

    # -- exact order, rank and maximum statistics ------------------------
    def _stage_pack(self, plans, c, ntrials):
        """Programs of `plans` and constants c (len(plans) * ntrials rows of
        kmax; None: each plan's start constants for every trial)."""
        words, off, pcount, c0p, kmax = pack_plans(plans, None)
        if c is None:
            c = np.repeat(c0p.reshape(len(plans), kmax), ntrials, axis=0)
        c = np.ascontiguousarray(np.asarray(c, np.float64).reshape(len(plans) * ntrials, -1))
        if c.shape[1] != kmax:
            cc = np.zeros((c.shape[0], kmax))
            w = min(kmax, c.shape[1])
            cc[:, :w] = c[:, :w]
            c = cc
        return words, off, pcount, np.ascontiguousarray(c.reshape(-1)), kmax

    def order_stat_batch(self, plans, what: int, c=None, ntrials: int = 1):
        """rc_order_stat_score_batch: (values, status, ms), one per segment."""
        words, off, pcount, c0p, kmax = self._stage_pack(plans, c, ntrials)
        ns = len(plans) * ntrials
        val = np.zeros(ns)
        st = np.zeros(ns, np.int32)
        ms = self._L.rc_order_stat_score_batch(
            self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p), _L.ptr(pcount, _L.i32p),
            len(plans), int(ntrials), kmax, _L.ptr(c0p, _L.f64p), int(what), _L.ptr(val, _L.f64p),
            _L.ptr(st, _L.i32p))
        return val, st, ms

    def rank_score_batch(self, plans, c=None, ntrials: int = 1):
        """rc_rank_score_batch: ([(sum AB, sum AA)], sum BB, status, ms)."""
        words, off, pcount, c0p, kmax = self._stage_pack(plans, c, ntrials)
        ns = len(plans) * ntrials
        out = np.zeros(4 * ns, np.uint64)
        bb = np.zeros(2, np.uint64)
        st = np.zeros(ns, np.int32)
        ms = self._L.rc_rank_score_batch(
            self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p), _L.ptr(pcount, _L.i32p),
            len(plans), int(ntrials), kmax, _L.ptr(c0p, _L.f64p), _L.ptr(out, _L.u64p),
            _L.ptr(st, _L.i32p), _L.ptr(bb, _L.u64p))
        sums = [(_i128(out[4 * s], out[4 * s + 1]), _i128(out[4 * s + 2], out[4 * s + 3]))
                for s in range(ns)]
        return sums, _i128(bb[0], bb[1]), st, ms

    def rank_debug_b(self) -> Optional[np.ndarray]:
        out = np.zeros(self.n, np.int64)
        rc = self._L.rc_rank_debug_b(self.ctx, out.ctypes.data_as(_L.C.POINTER(_L.C.c_longlong)))
        return None if rc != 0 else out

    def max_scan_batch(self, plans, c=None, ntrials: int = 1, coords=None, levels=None):
        """rc_max_scan_batch: ([RcMaxOut], ms)."""
        words, off, pcount, c0p, kmax = self._stage_pack(plans, c, ntrials)
        ns = len(plans) * ntrials
        out = (_L.RcMaxOut * ns)()
        co = None if coords is None else np.ascontiguousarray(coords, np.int32)
        lv = None if levels is None else np.ascontiguousarray(levels, np.float64)
        ms = self._L.rc_max_scan_batch(
            self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p), _L.ptr(pcount, _L.i32p),
            len(plans), int(ntrials), kmax, _L.ptr(c0p, _L.f64p), _L.ptr(co, _L.i32p),
            _L.ptr(lv, _L.f64p), out)
        return list(out), ms

    def fit_batch_hetero(self, plans, rowset, metric, iters: int, kmax=None):
        words, off, pcount, c0p, kmax = pack_plans(plans, kmax)
        n = len(plans)
        rs = np.ascontiguousarray(rowset, np.int32)
        mt = np.ascontiguousarray(metric, np.int32)
        outc = np.zeros((n, kmax))
        outsse = np.zeros(n)
        outst = np.zeros(n, np.int32)
        ms = self._L.rc_fit_batch_hetero(self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p),
                                         _L.ptr(pcount, _L.i32p), n, kmax, _L.ptr(c0p, _L.f64p),
                                         _L.ptr(rs, _L.i32p), _L.ptr(mt, _L.i32p), int(iters),
                                         _L.ptr(outc, _L.f64p), _L.ptr(outsse, _L.f64p),
                                         _L.ptr(outst, _L.i32p))
        return _fit_outputs(plans, outc, outsse, outst), ms

    def fit_jacobian(self, plans, c, cand: int, rows: int):
        words, off, pcount, c0p, kmax = pack_plans(plans)
        cc = np.ascontiguousarray(c, np.float64).reshape(len(plans), kmax)
        pred = np.zeros(rows, np.float32)
        jac = np.zeros((rows, kmax), np.float32)
        gx = np.zeros((rows, max(self.v, 1)), np.float32)
        rc = self._L.rc_fit_jacobian(self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p),
                                     _L.ptr(pcount, _L.i32p), len(plans), kmax,
                                     _L.ptr(cc, _L.f64p), int(cand), int(rows),
                                     _L.ptr(pred, _L.f32p), _L.ptr(jac, _L.f32p),
                                     _L.ptr(gx, _L.f32p))
        return (pred, jac, gx) if rc == 0 else None

    def implicit_score_batch(self, plans, metric: int = 0):
        words, off, pcount, c0p, kmax = pack_plans(plans)
        n = len(plans)
        out = np.zeros(n)
        st = np.zeros(n, np.int32)
        ms = self._L.rc_implicit_score_batch(self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p),
                                             _L.ptr(pcount, _L.i32p), n, kmax,
                                             _L.ptr(c0p, _L.f64p), int(metric),
                                             _L.ptr(out, _L.f64p), _L.ptr(st, _L.i32p))
        return out, st, ms

    def implicit_fit_batch(self, plans, iters: int):
        words, off, pcount, c0p, kmax = pack_plans(plans)
        n = len(plans)
        outc = np.zeros((n, kmax))
        outsse = np.zeros(n)
        outst = np.zeros(n, np.int32)
        ms = self._L.rc_implicit_fit_batch(self.ctx, _L.ptr(words, _L.u32p), _L.ptr(off, _L.u32p),
                                           _L.ptr(pcount, _L.i32p), n, kmax, _L.ptr(c0p, _L.f64p),
                                           int(iters), _L.ptr(outc, _L.f64p),
                                           _L.ptr(outsse, _L.f64p), _L.ptr(outst, _L.i32p))
        return _fit_outputs(plans, outc, outsse, outst), ms

    def stats(self) -> "_L.RcStats":
        s = _L.RcStats()
        self._L.rc_last_stats(self.ctx, C.byref(s))
        return s

    def arena_report(self) -> "_L.RcArenaReport":
        r = _L.RcArenaReport()
        self._L.rc_arena_report(self.ctx, C.byref(r))
        return r

    @staticmethod
    def device_name() -> str:
        name = _L.load().rc_device_name()
        return name.decode() if name else "unknown"


def probe_device_rates(rows: int = 2456306, fits: int = 64, programs: int = 256,
                       seed: int = 11):
    """The hardware probe behind dispatch.DeviceRates (dispatch.rs:642-660),
    measured on the current device with synthetic rows of three inputs:

      r_eval    row-evals/s of a batch of `programs` programs (21-31 nodes)
                over `rows` rows, from the batch's device time;
      r_fit     accepted fits/s at the reference shape (FIT_REF_ROWS rows,
                FIT_REF_CONSTANTS = 4 constants, FIT_REF_SWEEPS = 8 LM
                iterations): `fits` candidates in one batch, wall clock; a
                probe on fewer rows is scaled to FIT_REF_ROWS by the model's
                own inverse scaling in the rows;
      bw        host-to-device bytes/s of re-uploading the f64 fit rows;
      launch_s  the median wall time of a one-program batch over 64 rows.

    Raises RuntimeError when the device is missing or a call fails."""
    import time
    from . import dispatch
    from .expr import ADD, COS, MUL, SIN, const, node, var
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2.0, 2.0, size=(rows, 3))
    y = 1.5 * x[:, 0] - 0.7 * np.sin(x[:, 1]) + 0.3 * x[:, 0] * x[:, 2] + 0.2
    cols32 = np.ascontiguousarray(x.T, np.float32).reshape(-1)
    cols64 = np.ascontiguousarray(x.T).reshape(-1)
    ev = GpuEvaluator(cols32, rows, 3, y.astype(np.float32))
    try:
        # r_eval: programs of 21-31 nodes.
        pop = Population()
        while len(pop) < programs:
            t = var(int(rng.integers(0, 3)))
            for _ in range(int(rng.integers(5, 8))):
                u = node(SIN if rng.random() < 0.5 else COS, var(int(rng.integers(0, 3))))
                t = node(ADD, t, node(MUL, const(float(rng.normal())), u))
            pop.add(t)
        prog, off = pop.arrays()
        out, ms = ev.eval_words(prog, off, 0)
        out, ms = ev.eval_words(prog, off, 0)
        if out is None or ms <= 0:
            raise RuntimeError("evaluation probe failed")
        r_eval = len(pop) * rows / (ms * 1e-3)
        # bw: the second upload of the f64 rows reuses the device buffers.
        if not ev.set_fit_data(cols64, y):
            raise RuntimeError("set_fit_data failed")
        t0 = time.perf_counter()
        if not ev.set_fit_data(cols64, y):
            raise RuntimeError("set_fit_data failed")
        bw = (cols64.nbytes + y.nbytes) / (time.perf_counter() - t0)
        # r_fit: 4-constant candidates, 8 LM iterations.
        plans = []
        for _ in range(fits):
            c = rng.normal(size=4)
            t = node(ADD, node(ADD, node(MUL, const(c[0]), var(0)),
                               node(MUL, const(c[1]), node(SIN, var(1)))),
                     node(ADD, node(MUL, const(c[2]), node(MUL, var(0), var(2))), const(c[3])))
            plans.append(gpu_fit_plan(t))
        outs, ms = ev.fit_batch(plans, 8, 0)
        t0 = time.perf_counter()
        outs, ms = ev.fit_batch(plans, 8, 0)
        wall = time.perf_counter() - t0
        if ms < 0:
            raise RuntimeError("fit probe failed")
        r_fit = fits / wall * (rows / dispatch.FIT_REF_ROWS)
    finally:
        ev.close()
    small = GpuEvaluator(np.ascontiguousarray(x[:64].T, np.float32).reshape(-1), 64, 3,
                         y[:64].astype(np.float32))
    try:
        one = Population()
        one.add(node(ADD, var(0), var(1)))
        p1, o1 = one.arrays()
        walls = []
        for _ in range(21):
            t0 = time.perf_counter()
            out, ms = small.eval_words(p1, o1, 0)
            walls.append(time.perf_counter() - t0)
            if out is None or ms < 0:
                raise RuntimeError("launch probe failed")
        launch_s = float(np.median(walls[1:]))
    finally:
        small.close()
    return dispatch.DeviceRates(r_eval=r_eval, r_fit=r_fit, bw=bw, launch_s=launch_s)


class BatchStatus(Enum):
    """fit/gpu.rs:61-93."""
    GPU = "gpu"
    UNSUPPORTED = "unsupported"
    NON_FINITE = "non_finite"
    WORSE = "worse"


class BatchedGpu:
    """The BatchedFitter backed by the device (fit/gpu.rs:456-654).

    fit_batch returns, per tree, the proposed constants or None (declined,
    non-usable status, or a failed launch: the caller's exact arm refits).
    """

    def __init__(self, evaluator: GpuEvaluator):
        self.ev = evaluator

    def fit_batch_metric(self, trees, iters: int, metric: int) -> List[Optional[np.ndarray]]:
        plans = [gpu_fit_plan(t) for t in trees]
        idx = [i for i, p in enumerate(plans) if p is not None]
        out: List[Optional[np.ndarray]] = [None] * len(trees)
        if not idx:
            return out
        res, ms = self.ev.fit_batch([plans[i] for i in idx], iters, metric)
        if ms < 0:
            return out
        for j, i in enumerate(idx):
            if fit_status_usable(res[j].status) and np.all(np.isfinite(res[j].c)):
                out[i] = res[j].c
        return out


def fit_batch_verified(backend: BatchedGpu, trees, x, y, w, iters: int, metric: int, exact):
    """fit/gpu.rs:176-317: accept a proposal only if finite and no worse in
    exact f64 (SSE for squared, the metric otherwise); else the exact arm.

    Returns (trees, scores, statuses) with the trees carrying the kept
    constants.  `exact` supplies sse / metric_score / fit_metric.
    """
    proposals = backend.fit_batch_metric(trees, iters, metric)
    out_trees, scores, states = [], [], []
    for tree, c in zip(trees, proposals):
        if c is None:
            s, cc = exact.fit_metric(tree, x, y, w, iters, metric)
            out_trees.append(tree.with_constants(cc))
            scores.append(s)
            states.append(BatchStatus.UNSUPPORTED)
            continue
        trial = tree.with_constants(c)
        if metric == METRIC_SQUARED:
            start = exact.sse(tree, x, y, w)
            cand = exact.sse(trial, x, y, w)
        else:
            start = exact.metric_score(tree, x, y, w, metric)
            cand = exact.metric_score(trial, x, y, w, metric)
        if not np.isfinite(cand) or not np.isfinite(start) or cand > start:
            s, cc = exact.fit_metric(tree, x, y, w, iters, metric)
            out_trees.append(tree.with_constants(cc))
            scores.append(s)
            states.append(BatchStatus.WORSE)
        else:
            out_trees.append(trial)
            scores.append(cand)
            states.append(BatchStatus.GPU)
    return out_trees, scores, states
