"""rc_eval_static, the evaluation entry point that allocates nothing, against
rc_set_programs + rc_eval through the C ABI.

Two calls in a row (the second with a smaller batch) on a context with no
derived columns yet, then the same batches through the allocating path: the
losses are the same bits, the context's allocation counter does not move
during the static calls, and a batch larger than the static array is refused
with its own error code.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from rcsr_hip import host, lib, workloads

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _batch(trees):
    pop = host.Population()
    for t in trees:
        assert pop.add(t)
    return pop.arrays() + (len(pop),)


def _alloc_calls(ev):
    return int(ev.arena_report().alloc_calls)


def _static(ev, prog, off, nprog, loss):
    out = np.zeros(nprog, np.float32)
    ms = ev._L.rc_eval_static(ev.ctx, lib.ptr(prog, lib.u32p), lib.ptr(off, lib.u32p), nprog,
                              lib.ptr(out, lib.f32p), int(loss))
    return out, ms


def test_eval_static_matches_rc_eval_on_two_calls():
    wl = workloads.config2(n=50000, nprog=400)
    ev = host.GpuEvaluator(wl.cols32(), wl.n, wl.v, wl.y.astype(np.float32))
    a = _batch(wl.trees[:400])
    b = _batch(wl.trees[50:170])
    for name, (prog, off, nprog), loss in (("A", a, 0), ("B", b, 4)):
        _say(f"rc_eval_static batch {name}: {nprog} programs x n={wl.n}, loss {loss}")
        before = _alloc_calls(ev)
        got, ms = _static(ev, prog, off, nprog, loss)
        assert ms >= 0.0
        assert _alloc_calls(ev) == before
        want, ms2 = ev.eval_words(prog, off, loss)
        assert ms2 >= 0.0
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), name
    # Again after the allocating path has derived columns: still the same bits.
    prog, off, nprog = b
    before = _alloc_calls(ev)
    got, ms = _static(ev, prog, off, nprog, 4)
    assert ms >= 0.0 and _alloc_calls(ev) == before
    want, _ = ev.eval_words(prog, off, 4)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    ev.close()


def test_eval_static_refuses_a_batch_beyond_its_array():
    from rcsr_hip.expr import MUL, const, node, var
    n = 1 << 20
    x = np.linspace(-1.0, 1.0, n).astype(np.float32)
    ev = host.GpuEvaluator(x, n, 1, x)
    trees = [node(MUL, const(0.5 + 1e-4 * i), var(0)) for i in range(11000)]
    prog, off, nprog = _batch(trees)
    _say(f"rc_eval_static oversized batch: {nprog} programs x n={n}")
    before = _alloc_calls(ev)
    out, ms = _static(ev, prog, off, nprog, 0)
    assert ms == -2.0
    assert _alloc_calls(ev) == before
    # A batch that fits still runs on the same context.
    prog, off, nprog = _batch(trees[:100])
    out, ms = _static(ev, prog, off, nprog, 0)
    assert ms >= 0.0
    want, _ = ev.eval_words(prog, off, 0)
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))
    ev.close()
