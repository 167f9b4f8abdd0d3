"""The batched-evaluation workload (4096 programs of 3-60 nodes over 5 f32
input columns, N = 1 048 576) against the oracle, through the C ABI.

Mirrors the reference's evaluator checks (gpu.rs:3114-3210: batch equals
one-at-a-time, the loss sums equal an f64 reducer of the same f32
predictions).  At a reduced row count the device's f32 predictions of the
workload's own programs (the first 192 and 192 of 31-60 nodes, a batch the
evaluator's translation turns into derived columns and hoisted input-free
subtrees) are held against the f64 oracle; at the full row count the loss
sums of a sample of programs are held against an f64 reduction of the
device's own predictions of all 1 048 576 rows.
"""
import sys

import numpy as np
import pytest

import oracle
from rcsr_hip import host, lib, workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2():
    return workloads.config2()


def _size(t):
    return 1 + sum(_size(k) for k in t.kids) if t.kids else 1


def _translation_info(prog, off, nprog, nvars):
    import ctypes as C
    L = lib.load()
    info = np.zeros(8, np.int32)
    cap = 1 << 22
    code = np.zeros(cap, np.uint32)
    meta = np.zeros(cap, np.uint32)
    der = np.zeros(2 * 256, np.uint32)
    nder = C.c_int(0)
    rc = L.rch_eval_translate(lib.ptr(prog, lib.u32p), lib.ptr(off, lib.u32p), nprog, nvars,
                              lib.ptr(code, lib.u32p), cap, lib.ptr(meta, lib.u32p), cap,
                              lib.ptr(der, lib.u32p), 256, C.byref(nder), lib.ptr(info, lib.i32p))
    assert rc == 0
    nhoist = (int(info[1]) - 2 * nprog) // 2
    return nhoist, int(nder.value)


def test_c2_programs_match_the_oracle_at_reduced_rows(c2):
    n = 20000
    x = c2.x[:n]
    y = c2.y[:n]
    long_ones = [t for t in c2.trees if _size(t) >= 31][:192]
    trees = c2.trees[:192] + long_ones
    assert len(long_ones) == 192
    pop = host.Population()
    for t in trees:
        assert pop.add(t)
    prog, off = pop.arrays()
    nhoist, nder = _translation_info(prog, off, len(pop), 5)
    _say(f"C2 reduced: {len(trees)} programs (sizes {min(map(_size, trees))}-"
         f"{max(map(_size, trees))}), n={n}; {nhoist} hoisted subtrees, {nder} derived columns")
    assert nhoist > 0 and nder > 0
    ev = host.GpuEvaluator(c2.cols32(x), n, 5, y.astype(np.float32))
    pred = ev.predict_words(prog, off)
    assert pred is not None
    nan_mismatch = 0
    ok_rows = all_rows = 0
    med = []
    for t, p in zip(trees, pred):
        want = oracle.predict(t, x)
        nan_mismatch += int(np.sum(np.isnan(want) != np.isnan(p)))
        fin = np.isfinite(want) & (np.abs(want) < 1e6) & np.isfinite(p)
        if fin.sum() == 0:
            continue
        e = np.abs(p[fin].astype(np.float64) - want[fin]) / (1.0 + np.abs(want[fin]))
        med.append(np.median(e))
        ok_rows += int(np.sum(e < 1e-3))
        all_rows += int(fin.sum())
    med = np.asarray(med)
    _say(f"C2 reduced: median of per-program median error {np.median(med):.2e}, "
         f"worst per-program median {med.max():.2e}; rows within 1e-3: "
         f"{ok_rows / all_rows:.5f}; NaN mismatches {nan_mismatch} of {len(trees) * n}")
    # f32 rows and f32 operators against f64: long programs amplify rounding
    # near poles and guard boundaries, so the bound is on the typical row.
    assert np.median(med) < 1e-6
    assert np.all(med < 1e-3)
    assert ok_rows / all_rows > 0.99
    assert nan_mismatch <= 0.005 * len(trees) * n
    # Batch equals one-at-a-time for a sample of the programs.
    sums, ms = ev.eval_words(prog, off, 0)
    assert ms >= 0
    for i in (0, 7, 191, 200, 383):
        p1 = host.Population()
        p1.add(trees[i])
        g1, o1 = p1.arrays()
        one, _ = ev.eval_words(g1, o1, 0)
        assert one[0] == sums[i]
    ev.close()


def test_c2_full_rows_loss_sums_equal_an_f64_reduction(c2):
    n = c2.n
    rng = np.random.default_rng(17)
    pick = sorted(rng.choice(len(c2.trees), 16, replace=False).tolist())
    trees = [c2.trees[i] for i in pick]
    ev = host.GpuEvaluator(c2.cols32(), n, 5, c2.y.astype(np.float32))
    _say(f"C2 full rows: {len(trees)} sampled programs x n={n}")
    pop = host.Population()
    for t in trees:
        assert pop.add(t)
    prog, off = pop.arrays()
    sums, ms = ev.eval_words(prog, off, 0)
    assert ms >= 0
    preds = ev.predict_words(prog, off)
    y32 = c2.y.astype(np.float32)
    for i, t in enumerate(trees):
        p = preds[i]
        hard = ~np.isfinite(p) | (np.abs(p) > 1e30)
        r = (p - y32).astype(np.float32)
        want = float(np.sum(np.where(hard, 1e30, (r * r).astype(np.float64))))
        assert sums[i] == pytest.approx(np.float32(want), rel=1e-5), (pick[i], sums[i], want)
    ev.close()
