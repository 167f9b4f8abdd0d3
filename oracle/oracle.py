"""ctypes binding of the CPU f64 oracle — TEST INFRASTRUCTURE ONLY.

Loaded by the test suite, the smoke check and the benchmark's CPU-baseline
leg, always as the checker.  The product package never imports this module.
Trees are anything with a `.preorder()` method returning (ops, vals) arrays.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "librcsr_oracle.so")
_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.or_eval.restype = C.c_double
        L.or_eval.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int]
        L.or_erf.restype = C.c_double
        L.or_erf.argtypes = [C.c_double]
        L.or_erfc.restype = C.c_double
        L.or_erfc.argtypes = [C.c_double]
        L.or_nconst.restype = C.c_int
        L.or_nconst.argtypes = [_ip, _dp, C.c_int]
        L.or_sse.restype = C.c_double
        L.or_sse.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp]
        L.or_gradients.restype = None
        L.or_gradients.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, _dp, _dp]
        L.or_fit_weighted.restype = C.c_double
        L.or_fit_weighted.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                      C.c_int, _dp]
        L.or_fit_metric.restype = C.c_double
        L.or_fit_metric.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                    C.c_int, C.c_int, _dp]
        L.or_metric_score.restype = C.c_double
        L.or_metric_score.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                      C.c_int]
        L.or_implicit_score.restype = C.c_double
        L.or_implicit_score.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int]
        L.or_signed_distance.restype = C.c_int
        L.or_signed_distance.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, _dp]
        L.or_implicit_fit.restype = C.c_double
        L.or_implicit_fit.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, C.c_int,
                                      _dp]
        L.or_fit_batch.restype = None
        L.or_fit_batch.argtypes = [_ip, _dp, _ip, C.c_int, _ip, _dp, C.c_int, C.c_int, _dp,
                                   _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_implicit_fit_batch.restype = None
        L.or_implicit_fit_batch.argtypes = [_ip, _dp, _ip, C.c_int, _dp, C.c_int, C.c_int, _dp,
                                            C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.or_corr_fit.restype = C.c_double
        L.or_corr_fit.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int,
                                  C.c_int, C.c_int, C.c_int, _dp, _ip]
        L.or_corr_score.restype = C.c_double
        L.or_corr_score.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp, C.c_int]
        L.or_predict_flat.restype = None
        L.or_predict_flat.argtypes = [_ip, _dp, C.c_int, _dp, C.c_int, C.c_int, _dp]
        L.or_median_abs.restype = C.c_double
        L.or_median_abs.argtypes = [_dp, C.c_int]
        L.or_iqr_mean_abs.restype = C.c_double
        L.or_iqr_mean_abs.argtypes = [_dp, C.c_int]
        L.or_centred_ranks_doubled.restype = C.c_int
        L.or_centred_ranks_doubled.argtypes = [_dp, C.c_int, C.POINTER(C.c_longlong)]
        L.or_rank_sums.restype = None
        L.or_rank_sums.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.c_int,
                                   C.POINTER(C.c_ulonglong)]
        L.or_philox.restype = None
        L.or_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint32)]
        L.or_vary.restype = C.c_longlong
        L.or_vary.argtypes = [C.POINTER(C.c_uint32), C.c_int, C.c_int, C.c_int, _dp,
                              C.POINTER(C.c_ubyte), C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                              C.c_int, C.c_int, C.POINTER(C.c_uint32), C.c_int,
                              C.POINTER(C.c_uint32), C.POINTER(C.c_uint16),
                              C.POINTER(C.c_uint16), C.POINTER(C.c_uint64), _dp]
        L.or_max_scan.restype = None
        L.or_max_scan.argtypes = [_dp, _dp, C.c_int, C.c_double, _dp, _ip, _dp, _dp, _ip]
        _lib = L
    return _lib


def _i(a):
    return a.ctypes.data_as(_ip)


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _tree(t) -> Tuple[np.ndarray, np.ndarray]:
    ops, vals = t.preorder()
    return np.ascontiguousarray(ops, np.int32), np.ascontiguousarray(vals, np.float64)


def _rows(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    return np.ascontiguousarray(x)


def _w(w) -> Optional[np.ndarray]:
    if w is None or len(w) == 0:
        return None
    return np.ascontiguousarray(w, dtype=np.float64)


def eval_row(tree, x) -> float:
    ops, vals = _tree(tree)
    x = np.ascontiguousarray(x, np.float64)
    return lib().or_eval(_i(ops), _d(vals), len(ops), _d(x), len(x))


def predict(tree, X) -> np.ndarray:
    X = _rows(X)
    return np.array([eval_row(tree, X[i]) for i in range(X.shape[0])])


def erf(x: float) -> float:
    return lib().or_erf(float(x))


def erfc(x: float) -> float:
    return lib().or_erfc(float(x))


def nconst(tree) -> int:
    ops, vals = _tree(tree)
    return lib().or_nconst(_i(ops), _d(vals), len(ops))


def sse(tree, X, y, w=None) -> float:
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    return lib().or_sse(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(y), _d(ww))


def gradients(tree, x) -> Tuple[np.ndarray, np.ndarray]:
    ops, vals = _tree(tree)
    x = np.ascontiguousarray(x, np.float64)
    k = nconst(tree)
    g = np.zeros(max(k, 1))
    gx = np.zeros(max(len(x), 1))
    lib().or_gradients(_i(ops), _d(vals), len(ops), _d(x), len(x), _d(g), _d(gx))
    return g[:k], gx[:len(x)]


def fit_weighted(tree, X, y, w=None, iters=8) -> Tuple[float, np.ndarray]:
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    c = np.zeros(max(nconst(tree), 1))
    s = lib().or_fit_weighted(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(y),
                              _d(ww), int(iters), _d(c))
    return s, c[:nconst(tree)]


def fit_metric(tree, X, y, w=None, iters=8, metric=0) -> Tuple[float, np.ndarray]:
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    c = np.zeros(max(nconst(tree), 1))
    s = lib().or_fit_metric(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(y),
                            _d(ww), int(iters), int(metric), _d(c))
    return s, c[:nconst(tree)]


def metric_score(tree, X, y, w=None, metric=0) -> float:
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    return lib().or_metric_score(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1],
                                 _d(y), _d(ww), int(metric))


def implicit_score(tree, X, w=None, metric=0) -> float:
    ops, vals = _tree(tree)
    X = _rows(X)
    ww = _w(w)
    return lib().or_implicit_score(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1],
                                   _d(ww), int(metric))


def signed_distance(tree, x) -> Optional[float]:
    ops, vals = _tree(tree)
    x = np.ascontiguousarray(x, np.float64)
    d = np.zeros(1)
    rc = lib().or_signed_distance(_i(ops), _d(vals), len(ops), _d(x), len(x), _d(d))
    return None if rc != 0 else float(d[0])


def implicit_fit(tree, X, w=None, iters=8) -> Tuple[float, np.ndarray]:
    ops, vals = _tree(tree)
    X = _rows(X)
    ww = _w(w)
    k = nconst(tree)
    c = np.zeros(max(k, 1))
    s = lib().or_implicit_fit(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(ww),
                              int(iters), _d(c))
    return s, c[:k]


def fit_batch(trees: Sequence, X, y, w=None, iters=8, metrics=None, threads=1, kmax=20):
    """Fit many trees on the host (one tree per thread)."""
    ops_l, vals_l, off = [], [], [0]
    for t in trees:
        o, v = _tree(t)
        ops_l.append(o)
        vals_l.append(v)
        off.append(off[-1] + len(o))
    ops = np.ascontiguousarray(np.concatenate(ops_l), np.int32)
    vals = np.ascontiguousarray(np.concatenate(vals_l), np.float64)
    off = np.asarray(off, np.int32)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    met = None if metrics is None else np.ascontiguousarray(metrics, np.int32)
    n = len(trees)
    out_c = np.zeros((n, kmax))
    out_l = np.zeros(n)
    lib().or_fit_batch(_i(ops), _d(vals), _i(off), n, None if met is None else _i(met), _d(X),
                       X.shape[0], X.shape[1], _d(y), _d(ww), int(iters), int(kmax),
                       int(threads), _d(out_c), _d(out_l))
    return out_l, out_c


def implicit_fit_batch(trees: Sequence, X, w=None, iters=8, threads=1, kmax=8):
    ops_l, vals_l, off = [], [], [0]
    for t in trees:
        o, v = _tree(t)
        ops_l.append(o)
        vals_l.append(v)
        off.append(off[-1] + len(o))
    ops = np.ascontiguousarray(np.concatenate(ops_l), np.int32)
    vals = np.ascontiguousarray(np.concatenate(vals_l), np.float64)
    off = np.asarray(off, np.int32)
    X = _rows(X)
    ww = _w(w)
    n = len(trees)
    out_c = np.zeros((n, kmax))
    out_l = np.zeros(n)
    lib().or_implicit_fit_batch(_i(ops), _d(vals), _i(off), n, _d(X), X.shape[0], X.shape[1],
                                _d(ww), int(iters), int(kmax), int(threads), _d(out_c),
                                _d(out_l))
    return out_l, out_c


# -- order, rank and maximum statistics -----------------------------------

def median_abs(r) -> float:
    r = np.ascontiguousarray(r, np.float64)
    return float(lib().or_median_abs(_d(r), len(r)))


def iqr_mean_abs(r) -> float:
    r = np.ascontiguousarray(r, np.float64)
    return float(lib().or_iqr_mean_abs(_d(r), len(r)))


def centred_ranks_doubled(v) -> Optional[np.ndarray]:
    v = np.ascontiguousarray(v, np.float64)
    out = np.zeros(len(v), np.int64)
    rc = lib().or_centred_ranks_doubled(_d(v), len(v), out.ctypes.data_as(C.POINTER(C.c_longlong)))
    return None if rc != 0 else out


def _i128(hi, lo) -> int:
    u = (int(hi) << 64) | int(lo)
    return u - (1 << 128) if u >> 127 else u


def rank_sums(a, b) -> Tuple[int, int, int]:
    """Exact (sum a*b, sum a*a, sum b*b) of two centred doubled rank vectors."""
    a = np.ascontiguousarray(a, np.int64)
    b = np.ascontiguousarray(b, np.int64)
    out = np.zeros(6, np.uint64)
    lp = C.POINTER(C.c_longlong)
    lib().or_rank_sums(a.ctypes.data_as(lp), b.ctypes.data_as(lp), len(a),
                       out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
    return _i128(out[0], out[1]), _i128(out[2], out[3]), _i128(out[4], out[5])


def max_scan(r, b=None, level=float("nan")):
    """(vmax, row, lo, hi, status) of the exact maximum-error scan."""
    r = np.ascontiguousarray(r, np.float64)
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    vm, lo, hi = C.c_double(), C.c_double(), C.c_double()
    row, st = C.c_int(), C.c_int()
    lib().or_max_scan(_d(r), _d(bb), len(r), float(level), C.byref(vm), C.byref(row),
                      C.byref(lo), C.byref(hi), C.byref(st))
    return vm.value, row.value, lo.value, hi.value, st.value


def predict_flat(tree, X) -> np.ndarray:
    """Predictions of the flat (fit) form, the tree's own constants."""
    ops, vals = _tree(tree)
    X = _rows(X)
    out = np.zeros(X.shape[0])
    lib().or_predict_flat(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(out))
    return out


def corr_fit(tree, X, y, w=None, iters=8, metric=2, gauge=(-1, -1)):
    """The device correlation arm restated: (objective, constants, status)."""
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    ww = _w(w)
    out = np.zeros(max(nconst(tree), 1))
    st = C.c_int()
    obj = lib().or_corr_fit(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(y),
                            _d(ww), int(iters), int(metric), int(gauge[0]), int(gauge[1]),
                            _d(out), C.byref(st))
    return obj, out[:nconst(tree)], st.value


def corr_score(tree, X, y, w=None, metric=2) -> float:
    ops, vals = _tree(tree)
    X = _rows(X)
    y = np.ascontiguousarray(y, np.float64)
    return lib().or_corr_score(_i(ops), _d(vals), len(ops), _d(X), X.shape[0], X.shape[1], _d(y),
                               _d(_w(w)), int(metric))


def philox(key, ctr):
    """Philox4x32-10 of one (key, counter) block (genome.rs:214-300)."""
    k = np.ascontiguousarray(key, np.uint32)
    c = np.ascontiguousarray(ctr, np.uint32)
    out = np.zeros(4, np.uint32)
    u32 = C.POINTER(C.c_uint32)
    lib().or_philox(k.ctypes.data_as(u32), c.ctypes.data_as(u32), out.ctypes.data_as(u32))
    return [int(v) for v in out]


# The reference's default operator set (expr.rs:392-398) as (opcode, arity)
# pairs in table order: neg exp ln abs sqrt sin cos tan sinh cosh tanh
# logistic gauss square cube add sub mul div pow if.
DEFAULT_OPS = [(17, 1), (9, 1), (10, 1), (12, 1), (11, 1), (6, 1), (7, 1), (8, 1), (21, 1),
               (22, 1), (13, 1), (14, 1), (18, 1), (15, 1), (16, 1), (2, 2), (3, 2), (4, 2),
               (5, 2), (20, 2), (32, 3)]


def vary(parents, islands, pop, err, seed, gen, max_complexity=30, nvars=2, init_depth=3,
         dedup=True, restart=None, ops=None):
    """One generation of the flat variation engine (postfix.rs:500-893)."""
    parents = np.ascontiguousarray(parents, np.uint32)
    n, width = parents.shape
    assert n == islands * pop
    err = np.ascontiguousarray(err, np.float64)
    rs = None if restart is None else np.ascontiguousarray(restart, np.uint8)
    table = np.ascontiguousarray(np.asarray(ops or DEFAULT_OPS, np.uint32).reshape(-1))
    words = np.zeros((n, width), np.uint32)
    ln = np.zeros(n, np.uint16)
    cost = np.zeros(n, np.uint16)
    h = np.zeros(n, np.uint64)
    e = np.zeros(n, np.float64)
    u32 = C.POINTER(C.c_uint32)
    att = lib().or_vary(parents.ctypes.data_as(u32), islands, pop, width, _d(err),
                        None if rs is None else rs.ctypes.data_as(C.POINTER(C.c_ubyte)),
                        seed, gen, max_complexity, nvars, init_depth, int(dedup),
                        table.ctypes.data_as(u32), len(table) // 2, words.ctypes.data_as(u32),
                        ln.ctypes.data_as(C.POINTER(C.c_uint16)),
                        cost.ctypes.data_as(C.POINTER(C.c_uint16)),
                        h.ctypes.data_as(C.POINTER(C.c_uint64)), _d(e))
    assert att >= 0
    return {"words": words, "len": ln, "cost": cost, "hash": h, "err": e, "attempts": int(att)}
