"""ctypes binding of librcsrhip.so (the C ABI in include/rcsr_hip.h).

This is the reference-side binding a Python caller adds; the Rust binding of
the same symbols is gpu.rs:668-825.  Loading fails loudly when the library is
missing: there is no CPU fallback behind any of these calls.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_DIR, "librcsrhip.so")

u32p = C.POINTER(C.c_uint32)
i32p = C.POINTER(C.c_int)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_ulonglong)
u8p = C.POINTER(C.c_ubyte)
u16p = C.POINTER(C.c_uint16)
vp = C.c_void_p


class RcStats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("launch_ms", C.c_double), ("sync_ms", C.c_double),
                ("alloc_ms", C.c_double), ("h2d_ms", C.c_double), ("d2h_ms", C.c_double),
                ("wall_ms", C.c_double), ("h2d_bytes", C.c_ulonglong),
                ("d2h_bytes", C.c_ulonglong), ("batch", C.c_int), ("failed", C.c_int)]


class RcArenaReport(C.Structure):
    _fields_ = [("async_copy", C.c_int), ("reason", C.c_char * 96),
                ("pinned_bytes", C.c_ulonglong), ("alloc_calls", C.c_ulonglong),
                ("steady_alloc_calls", C.c_ulonglong), ("overlapped_uploads", C.c_ulonglong),
                ("batches", C.c_ulonglong), ("fs_alloc_calls", C.c_ulonglong),
                ("fs_growths", C.c_ulonglong), ("fs_steady_alloc_calls", C.c_ulonglong),
                ("fs_batches", C.c_ulonglong), ("fs_dev_bytes", C.c_ulonglong),
                ("fs_pinned_bytes", C.c_ulonglong)]


class RcMaxOut(C.Structure):
    _fields_ = [("vmax", C.c_double), ("lo", C.c_double), ("hi", C.c_double), ("row", C.c_int),
                ("status", C.c_int)]


class RcFitReport(C.Structure):
    _fields_ = [("asm_ms", C.c_double), ("trial_ms", C.c_double), ("exact_ms", C.c_double),
                ("ctrl_ms", C.c_double), ("asm_launches", C.c_int),
                ("trial_launches", C.c_int), ("exact_launches", C.c_int),
                ("ctrl_launches", C.c_int)]


# name -> (restype, argtypes); the complete exported surface of rcsr_hip.h
# and rcsr_host.h.
SIGNATURES = {
    "rc_create": (vp, [f32p, C.c_int, C.c_int, f32p]),
    "rc_destroy": (None, [vp]),
    "rc_set_fit_data": (C.c_int, [vp, f64p, f64p, f64p]),
    "rc_set_score_weights": (C.c_int, [vp, f64p]),
    "rc_set_fit_sign": (C.c_int, [vp, f64p]),
    "rc_set_programs": (C.c_int, [vp, u32p, u32p, C.c_int]),
    "rc_eval_submit": (C.c_int, [vp, C.c_int]),
    "rc_eval_collect": (C.c_float, [vp, f32p]),
    "rc_eval": (C.c_float, [vp, f32p, C.c_int]),
    "rc_predict": (C.c_int, [vp, u32p, u32p, C.c_int, f32p]),
    "rc_score_batch": (C.c_float, [vp, u32p, u32p, C.c_int, C.c_int, i32p, f64p, i32p]),
    "rc_pool_reserve": (C.c_int, [vp, C.c_int, C.c_int]),
    "rc_pool_upload": (C.c_int, [vp, C.c_int, u32p, C.c_int]),
    "rc_pool_eval": (C.c_float, [vp, C.c_int, C.c_int, C.c_int, f32p, C.c_int]),
    "rc_pool_score": (C.c_float, [vp, C.c_int, C.c_int, C.c_int, C.c_int, i32p, f64p, i32p]),
    "rc_pool_hash": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, u64p]),
    "rc_pool_vary": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_uint, C.c_uint, C.c_int, C.c_int,
                               C.c_int, C.c_int, u32p, C.c_int, f64p, u8p, u16p, u16p, u64p,
                               f64p]),
    "rc_pool_download": (C.c_int, [vp, C.c_int, u32p, C.c_int]),
    "rc_fit_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, f64p, i32p, C.c_int,
                                 C.c_int, f64p, f64p, i32p]),
    "rc_fit_kernel_info": (C.c_int, [C.c_int, i32p, i32p, i32p]),
    "rc_set_fit_mode": (C.c_int, [vp, C.c_int]),
    "rc_set_tuning": (C.c_int, [vp, C.c_int, C.c_int]),
    "rc_get_tuning": (C.c_int, [vp, C.c_int]),
    "rc_add_rows": (C.c_int, [vp, f32p, C.c_int, C.c_int, f32p, f64p, f64p, f64p]),
    "rc_fit_batch_hetero": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, f64p, i32p, i32p,
                                        C.c_int, f64p, f64p, i32p]),
    "rc_fit_jacobian": (C.c_int, [vp, u32p, u32p, i32p, C.c_int, C.c_int, f64p, C.c_int, C.c_int,
                                  f32p, f32p, f32p]),
    "rc_implicit_score_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, f64p, C.c_int,
                                            f64p, i32p]),
    "rc_implicit_fit_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, f64p, C.c_int,
                                          f64p, f64p, i32p]),
    "rc_fit_profile": (C.c_int, [vp, C.c_int]),
    "rc_fit_report": (C.c_int, [vp, C.POINTER(RcFitReport), i32p, C.c_int]),
    "rc_stats_begin": (None, [vp]),
    "rc_last_stats": (C.c_int, [vp, C.POINTER(RcStats)]),
    "rc_arena_report": (C.c_int, [vp, C.POINTER(RcArenaReport)]),
    "rc_arena_force_sync": (C.c_int, [vp, C.c_int]),
    "rc_device_count": (C.c_int, []),
    "rc_device_name": (C.c_char_p, []),
    "rc_unified": (C.c_int, []),
    "rch_compile_program": (C.c_int, [i32p, f64p, C.c_int, u32p, C.c_int, i32p]),
    "rch_fit_plan": (C.c_int, [i32p, f64p, C.c_int, u32p, C.c_int, f64p, C.c_int, i32p]),
    "rch_fitvm_compile": (C.c_int, [u32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint, u32p,
                                    C.c_int, u32p, C.c_int, i32p]),
    "rc_order_stat_score_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, C.c_int,
                                              f64p, C.c_int, f64p, i32p]),
    "rc_rank_score_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, C.c_int, f64p,
                                        u64p, i32p, u64p]),
    "rc_rank_debug_b": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
    "rc_max_scan_batch": (C.c_float, [vp, u32p, u32p, i32p, C.c_int, C.c_int, C.c_int, f64p,
                                      i32p, f64p, C.POINTER(RcMaxOut)]),
    "rch_eval_translate": (C.c_int, [u32p, u32p, C.c_int, C.c_int, u32p, C.c_int, u32p, C.c_int,
                                     u32p, C.c_int, i32p, i32p]),
    "rc_eval_static": (C.c_float, [vp, u32p, u32p, C.c_int, f32p, C.c_int]),
    "rch_fit_graph_key": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_int, u64p, C.c_int]),
}

_lib = None


def load():
    """Load librcsrhip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"librcsrhip.so not found at {LIB_PATH}: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a, t):
    """ctypes pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(t)
