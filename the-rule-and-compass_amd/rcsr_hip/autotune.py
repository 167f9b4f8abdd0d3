"""Autotuner of the device arm's launch shapes.

The shapes (rc_set_tuning, include/rcsr_hip.h):

  - rows per fit work item (TUNE_FIT_TILE_ROWS): how finely a candidate's
    rows are cut into single-wave work items.  Small tiles give small or
    few-candidate batches enough waves; large tiles cut the per-item partial
    Grams the control step reduces, which matters for candidates with
    millions of rows;
  - rows per lane of the fit VM assembly (TUNE_FIT_ROWS_PER_LANE): 4 rows
    share each interpreted instruction, 2 rows halve the register files and
    the LDS block (more waves per SIMD);
  - rows per evaluation tile (TUNE_EVAL_TILE_ROWS).

The tuner times the caller's own work (a fit batch, a sample of it, or an
evaluation batch) under each shape of a small grid on the device, keeps the
fastest, and leaves the built-in shape in place unless another is faster by
more than `margin` (timing noise never flips a choice).  Choices can be kept
in a JSON cache keyed by a workload signature (row counts, batch size,
constants, metrics in power-of-two buckets), so a search loop tunes once per
shape of work.  A candidate's results depend on its program, its rows and
the shapes, never on the rest of its batch; two shapes may differ in the
last bits of the fitted losses (the per-item partial sums are grouped
differently).
"""
from __future__ import annotations

import json
import math
import os
import tempfile
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from .host import TUNE_EVAL_TILE_ROWS, TUNE_FIT_ROWS_PER_LANE, TUNE_FIT_TILE_ROWS

FIT_TILES = (2048, 4096, 8192, 16384, 32768, 65536)
FIT_ROWS = (4, 2)
EVAL_TILES = (1024, 2048, 4096, 8192, 16384)
BUILTIN = {TUNE_FIT_TILE_ROWS: 8192, TUNE_FIT_ROWS_PER_LANE: 4, TUNE_EVAL_TILE_ROWS: 4096}
BUILTIN_FIT = {TUNE_FIT_TILE_ROWS: 8192, TUNE_FIT_ROWS_PER_LANE: 4}

Setting = Dict[int, int]


@dataclass
class TuneResult:
    best: Setting
    best_ms: float
    builtin_ms: float
    table: List[Tuple[Setting, float]] = field(default_factory=list)
    # With tune_fit(confirm=True): the sample winner and the built-in shape
    # timed on the whole batch (empty when the sample kept the built-in).
    confirm_table: List[Tuple[Setting, float]] = field(default_factory=list)

    @property
    def speedup(self) -> float:
        return self.builtin_ms / self.best_ms if self.best_ms > 0 else float("nan")

    def to_json(self) -> dict:
        name = {TUNE_FIT_TILE_ROWS: "fit_tile_rows", TUNE_FIT_ROWS_PER_LANE: "fit_rows_per_lane",
                TUNE_EVAL_TILE_ROWS: "eval_tile_rows"}
        return {
            "best": {name[k]: v for k, v in self.best.items()},
            "best_ms": self.best_ms,
            "builtin_ms": self.builtin_ms,
            "table": [[{name[k]: v for k, v in s.items()}, ms] for s, ms in self.table],
            "confirm_table": [[{name[k]: v for k, v in s.items()}, ms]
                              for s, ms in self.confirm_table],
        }


def _is_builtin(setting: Setting) -> bool:
    return all(BUILTIN.get(k) == v for k, v in setting.items())


def tune(ev, run: Callable[[], float], space: Sequence[Setting], reps: int = 2,
         margin: float = 0.02) -> Optional[TuneResult]:
    """Time run() (device ms of one call, < 0 = failed) under every setting of
    `space` on evaluator `ev`, apply the fastest (the built-in setting unless
    another beats it by more than `margin`) and return the timings.  Settings
    the library refuses, and runs that fail, are skipped; with no successful
    run the previous shapes are restored and None is returned."""
    keys = sorted({k for s in space for k in s})
    saved = {k: ev.get_tuning(k) for k in keys}
    table: List[Tuple[Setting, float]] = []
    for setting in space:
        if any(ev.set_tuning(k, v) < 0 for k, v in setting.items()):
            continue
        times = [run() for _ in range(max(reps, 1))]
        good = [t for t in times if t >= 0]
        if good:
            table.append((dict(setting), float(min(good))))
    if not table:
        for k, v in saved.items():
            ev.set_tuning(k, v)
        return None
    best, best_ms = min(table, key=lambda e: e[1])
    builtin = [ms for s, ms in table if _is_builtin(s)]
    builtin_ms = builtin[0] if builtin else float("nan")
    if builtin and builtin_ms <= best_ms * (1.0 + margin):
        best, best_ms = next(s for s, ms in table if _is_builtin(s)), builtin_ms
    for k, v in best.items():
        ev.set_tuning(k, v)
    return TuneResult(dict(best), best_ms, builtin_ms, table)


def fit_space(max_rows: int, tiles: Sequence[int] = FIT_TILES,
              rows: Sequence[int] = FIT_ROWS) -> List[Setting]:
    """Tile sizes up to the next one covering the largest candidate (larger
    tiles are the same work items), crossed with the rows-per-lane choices."""
    cover = [t for t in tiles if t < max_rows]
    bigger = [t for t in tiles if t >= max_rows]
    if bigger:
        cover.append(min(bigger))
    if BUILTIN[TUNE_FIT_TILE_ROWS] not in cover and BUILTIN[TUNE_FIT_TILE_ROWS] in tiles:
        cover.append(BUILTIN[TUNE_FIT_TILE_ROWS])
    return [{TUNE_FIT_TILE_ROWS: t, TUNE_FIT_ROWS_PER_LANE: r} for t in sorted(set(cover))
            for r in rows]


def eval_space(n: int, tiles: Sequence[int] = EVAL_TILES) -> List[Setting]:
    cover = [t for t in tiles if t < n] + [t for t in tiles if t >= n][:1]
    if BUILTIN[TUNE_EVAL_TILE_ROWS] not in cover and BUILTIN[TUNE_EVAL_TILE_ROWS] in tiles:
        cover.append(BUILTIN[TUNE_EVAL_TILE_ROWS])
    return [{TUNE_EVAL_TILE_ROWS: t} for t in sorted(set(cover))]


def _b(x: float) -> int:
    return int(math.floor(math.log2(max(float(x), 1.0))))


def signature(kind: str, rows: Sequence[int], ks: Sequence[int] = (),
              metrics: Sequence[int] = ()) -> str:
    """Power-of-two buckets of what sets the best shape: the batch size, the
    median and largest row counts, the mean constant count and the metric
    mix."""
    rows = sorted(int(r) for r in rows)
    med = rows[len(rows) // 2] if rows else 0
    mx = rows[-1] if rows else 0
    km = sum(ks) / len(ks) if ks else 0.0
    mix = "none"
    if metrics:
        l1 = sum(1 for m in metrics if m == 4) / len(metrics)
        mix = "sq" if l1 == 0 else ("l1" if l1 == 1 else "mixed")
    return f"{kind}:P{_b(len(rows))}:med{_b(med)}:max{_b(mx)}:k{_b(km + 1)}:{mix}"


class TuningCache:
    """Chosen shapes per workload signature in a JSON file."""

    def __init__(self, path: str):
        self.path = path
        self.data: Dict[str, Dict[str, int]] = {}
        if os.path.exists(path):
            with open(path) as fh:
                raw = json.load(fh)
            self.data = {k: {str(kk): int(vv) for kk, vv in v.items()} for k, v in raw.items()}

    def get(self, sig: str) -> Optional[Setting]:
        v = self.data.get(sig)
        return None if v is None else {int(k): int(x) for k, x in v.items()}

    def put(self, sig: str, setting: Setting) -> None:
        self.data[sig] = {str(k): int(v) for k, v in setting.items()}
        d = os.path.dirname(os.path.abspath(self.path))
        os.makedirs(d, exist_ok=True)
        fd, tmp = tempfile.mkstemp(dir=d, prefix=".tuning.")
        with os.fdopen(fd, "w") as fh:
            json.dump(self.data, fh, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def _apply(ev, setting: Setting) -> None:
    for k, v in setting.items():
        ev.set_tuning(k, v)


def tune_fit(ev, plans, iters: int, metric: int = 0, rows: Optional[Sequence[int]] = None,
             sample: int = 256, reps: int = 2, cache: Optional[TuningCache] = None,
             rowset=None, metrics=None, space: Optional[Sequence[Setting]] = None,
             confirm: bool = False, margin: float = 0.02) -> Optional[TuneResult]:
    """Tune the fit shapes for a batch: `rows` are the candidates' row counts
    (one per candidate); a heterogeneous batch passes its rowset and metrics
    (rc_fit_batch_hetero).  The timing runs on an evenly spread sample of at
    most `sample` candidates.  With `confirm`, a sample winner other than the
    built-in shape is timed against the built-in shape on the whole batch and
    kept only if it is faster there by more than `margin` (a sample can rank
    shapes differently from the batch it stands for).  With a cache, a known
    signature is applied without timing (best_ms is then NaN)."""
    n = len(plans)
    rows = list(rows) if rows is not None else [0] * n
    step = max(1, n // max(sample, 1))
    idx = list(range(0, n, step))[:sample]
    ks = [plans[i].nconst for i in idx]
    mets = [int(metrics[i]) for i in idx] if metrics is not None else [metric] * len(idx)
    sig = signature("fit", [rows[i] for i in idx], ks, mets)
    if cache is not None:
        known = cache.get(sig)
        if known is not None:
            _apply(ev, known)
            return TuneResult(known, float("nan"), float("nan"))
    sub = [plans[i] for i in idx]
    if rowset is not None:
        rs = [int(rowset[i]) for i in idx]

        def run():
            return ev.fit_batch_hetero(sub, rs, mets, iters)[1]
    else:
        def run():
            return ev.fit_batch(sub, iters, metric)[1]
    res = tune(ev, run, space if space is not None else fit_space(max(rows) if rows else 0),
               reps=reps, margin=margin)
    if res is not None and confirm and not _is_builtin(res.best):
        if rowset is not None:
            rs_all = [int(r) for r in rowset]
            ms_all = [int(m) for m in metrics] if metrics is not None else [metric] * n

            def run_all():
                return ev.fit_batch_hetero(plans, rs_all, ms_all, iters)[1]
        else:
            def run_all():
                return ev.fit_batch(plans, iters, metric)[1]
        full = tune(ev, run_all, [dict(BUILTIN_FIT), dict(res.best)], reps=reps, margin=margin)
        if full is not None:
            res = TuneResult(full.best, full.best_ms, full.builtin_ms, res.table, full.table)
    if res is not None and cache is not None:
        cache.put(sig, res.best)
    return res


def tune_eval(ev, prog, off, n: int, reps: int = 3, cache: Optional[TuningCache] = None,
              space: Optional[Sequence[Setting]] = None) -> Optional[TuneResult]:
    """Tune the evaluation tile for a program batch over n rows."""
    nprog = len(off) - 1
    sig = signature("eval", [n] * nprog)
    if cache is not None:
        known = cache.get(sig)
        if known is not None:
            _apply(ev, known)
            return TuneResult(known, float("nan"), float("nan"))

    def run():
        return ev.eval_words(prog, off, 0)[1]
    res = tune(ev, run, space if space is not None else eval_space(n), reps=reps)
    if res is not None and cache is not None:
        cache.put(sig, res.best)
    return res
