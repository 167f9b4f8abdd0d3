"""Device-side variation of the resident genome pool (rc_pool_vary) against
the oracle's restatement of the flat postfix engine (postfix.rs:500-893)
and Philox4x32-10 (genome.rs:214-300), through the C ABI.

The parents live in pool arena 0; one call writes the children into arena
1 on the device.  Every child word, length, node count, hash and error must
be the oracle's bit for bit, with and without dedup, with a restarted
island, over two chained generations (the second scored by the device
evaluator), and the pool hash kernel must return the children's hashes."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from rcsr_hip import host, lib

pytestmark = pytest.mark.gpu

HALT = 255


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _parents(islands, pop, width, seed):
    blank = np.full((islands * pop, width), HALT, np.uint32)
    blank[:, 0], blank[:, 1], blank[:, 2] = 1, 0, HALT
    g = oracle.vary(blank, islands, pop, np.full(islands * pop, np.nan), seed, 0, nvars=2,
                    restart=[1] * islands)
    err = np.array([abs(np.sin(0.37 * i + 0.1)) for i in range(islands * pop)])
    err[::11] = np.nan
    return g["words"], err


def _device_vary(ev, parents, err, islands, pop, seed, gen, dedup, restart=None, maxc=30,
                 nvars=2, depth=3):
    L = ev._L
    n, width = parents.shape
    assert L.rc_pool_reserve(ev.ctx, n, width) == 0
    words = np.ascontiguousarray(parents, np.uint32)
    assert L.rc_pool_upload(ev.ctx, 0, lib.ptr(words.reshape(-1), lib.u32p), n) == 0
    ops = np.ascontiguousarray(np.asarray(oracle.DEFAULT_OPS, np.uint32).reshape(-1))
    rs = None if restart is None else np.ascontiguousarray(restart, np.uint8)
    ln = np.zeros(n, np.uint16)
    cost = np.zeros(n, np.uint16)
    hsh = np.zeros(n, np.uint64)
    e = np.zeros(n)
    errc = np.ascontiguousarray(err, np.float64)
    rc = L.rc_pool_vary(ev.ctx, 0, islands, pop, seed, gen, maxc, nvars, depth, int(dedup),
                        lib.ptr(ops, lib.u32p), len(ops) // 2, lib.ptr(errc, lib.f64p),
                        None if rs is None else rs.ctypes.data_as(C.POINTER(C.c_ubyte)),
                        ln.ctypes.data_as(C.POINTER(C.c_uint16)),
                        cost.ctypes.data_as(C.POINTER(C.c_uint16)),
                        lib.ptr(hsh, lib.u64p), lib.ptr(e, lib.f64p))
    assert rc == 0
    out = np.zeros(n * width, np.uint32)
    assert L.rc_pool_download(ev.ctx, 1, lib.ptr(out, lib.u32p), n) == 0
    return {"words": out.reshape(n, width), "len": ln, "cost": cost, "hash": hsh, "err": e}


def _same(dev, ref):
    assert np.array_equal(dev["words"], ref["words"])
    assert np.array_equal(dev["len"], ref["len"])
    assert np.array_equal(dev["cost"], ref["cost"])
    assert np.array_equal(dev["hash"], ref["hash"])
    assert np.array_equal(dev["err"].view(np.uint64), ref["err"].view(np.uint64))


@pytest.fixture(scope="module")
def ev():
    rng = np.random.default_rng(2)
    n = 2048
    x = rng.uniform(-2, 2, size=(n, 2))
    y = np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2
    e = host.GpuEvaluator(x.T.astype(np.float32).reshape(-1), n, 2, y.astype(np.float32))
    yield e
    e.close()


@pytest.mark.parametrize("dedup", [True, False])
def test_device_variation_matches_the_oracle(ev, dedup):
    islands, pop, width = 4, 96, 64
    parents, err = _parents(islands, pop, width, seed=21)
    _say(f"pool vary: {islands} x {pop}, width {width}, dedup {dedup}")
    dev = _device_vary(ev, parents, err, islands, pop, 21, 1, dedup)
    ref = oracle.vary(parents, islands, pop, err, 21, 1, nvars=2, dedup=dedup)
    _same(dev, ref)
    # The pool hash kernel agrees with the children's FNV-1a hashes.
    hashes = np.zeros(islands * pop, np.uint64)
    assert ev._L.rc_pool_hash(ev.ctx, 1, islands * pop, width, lib.ptr(hashes, lib.u64p)) == 0
    assert np.array_equal(hashes, ref["hash"])


def test_device_variation_restart_and_a_second_generation(ev):
    islands, pop, width = 3, 64, 80
    parents, err = _parents(islands, pop, width, seed=5)
    dev = _device_vary(ev, parents, err, islands, pop, 5, 1, True, restart=[0, 1, 0], maxc=24,
                       depth=4)
    ref = oracle.vary(parents, islands, pop, err, 5, 1, max_complexity=24, nvars=2,
                      init_depth=4, restart=[0, 1, 0])
    _same(dev, ref)
    # Score the children on the device (the pool's own evaluation) and breed
    # the next generation from them.
    n = islands * pop
    sums = np.zeros(n, np.float32)
    assert ev._L.rc_pool_eval(ev.ctx, 1, n, width, lib.ptr(sums, lib.f32p), 0) >= 0
    err2 = sums.astype(np.float64) / ev.n
    err2[~np.isfinite(err2)] = np.nan
    dev2 = _device_vary(ev, dev["words"], err2, islands, pop, 5, 2, True, maxc=24, depth=4)
    ref2 = oracle.vary(dev["words"], islands, pop, err2, 5, 2, max_complexity=24, nvars=2,
                       init_depth=4)
    _same(dev2, ref2)
    _say(f"pool vary: two generations of {n}, {len(set(ref2['hash'].tolist()))} distinct")


@pytest.mark.parametrize("width,islands,pop,maxc,depth", [(16, 2, 5, 6, 2), (128, 1, 40, 60, 6)])
def test_device_variation_at_the_slot_width_limits(ev, width, islands, pop, maxc, depth):
    # A narrow pool, where most attempts fail the width filter and fall back
    # to random programs, and the widest supported slot with deep subtrees.
    parents, err = _parents(islands, pop, width, seed=13)
    _say(f"pool vary: {islands} x {pop}, width {width}, max complexity {maxc}, depth {depth}")
    for dedup in (True, False):
        dev = _device_vary(ev, parents, err, islands, pop, 13, 3, dedup, maxc=maxc, depth=depth)
        ref = oracle.vary(parents, islands, pop, err, 13, 3, max_complexity=maxc, nvars=2,
                          init_depth=depth, dedup=dedup)
        _same(dev, ref)
