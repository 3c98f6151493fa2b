"""ctypes binding of libgapx.so (include/gapx.h).

The product path has no CPU fallback: if the HIP library is missing or cannot
load, every entry point raises GapxError.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgapx.so")

WANT_E, WANT_F, WANT_V, WANT_LOCAL_E, WANT_LOCAL_V, WANT_VARIANCE = 1, 2, 4, 8, 16, 32
PREC = {"fp64": 0, "bf16x3": 1, "fp16x3": 2}


class GapxError(RuntimeError):
    """A non-zero GAPX_E_* status; `code` holds it."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


_lib = None

# (name, restype, argtypes); every int-returning call is status-checked
_SIGS = [
    ("gapx_api_version", C.c_int, []),
    ("gapx_last_error", None, [C.c_char_p, C.c_int]),
    ("gapx_device_count", C.c_int, [C.POINTER(C.c_int)]),
    ("gapx_model_load", C.c_int, [C.c_char_p, C.c_int64, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_void_p)]),
    ("gapx_model_free", C.c_int, [C.c_void_p]),
    ("gapx_model_cutoff", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("gapx_model_set_e_scale", C.c_int, [C.c_void_p, C.c_double]),
    ("gapx_model_set_precision", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_model_info", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    ("gapx_calc", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                            C.c_uint] + [C.c_void_p] * 5),
    ("gapx_calc_ex", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                               C.c_uint, C.c_void_p, C.c_int, C.c_double] + [C.c_void_p] * 8),
    ("gapx_calc_device_ex", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_double, C.c_uint, C.c_void_p, C.c_int, C.c_double] + [C.c_void_p] * 8),
    ("gapx_descriptor_variances", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                            C.POINTER(C.c_int)]),
    ("gapx_calc_lammps", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
     + [C.c_void_p] * 10),
    ("gapx_calc_device", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                                   C.c_uint] + [C.c_void_p] * 5),
    ("gapx_sync", C.c_int, [C.c_void_p]),
    ("gapx_calc_device_noalloc", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_double, C.c_uint] + [C.c_void_p] * 5),
    ("gapx_calc_batch", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_double, C.c_uint] + [C.c_void_p] * 5),
    ("gapx_calc_batch_device", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_double, C.c_uint] + [C.c_void_p] * 5),
    ("gapx_server_create", C.c_int, [C.c_void_p, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_void_p)]),
    ("gapx_server_submit", C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint,
                                     C.POINTER(C.c_int64)]),
    ("gapx_server_poll", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int)]),
    ("gapx_server_wait", C.c_int, [C.c_void_p, C.c_int64] + [C.c_void_p] * 5 + [C.POINTER(C.c_double)]),
    ("gapx_server_stats", C.c_int, [C.c_void_p] + [C.POINTER(C.c_int64)] * 3 + [C.POINTER(C.c_double)]),
    ("gapx_server_free", C.c_int, [C.c_void_p]),
    ("gapx_server_session_open", C.c_int, [C.c_void_p, C.c_double, C.POINTER(C.c_int64)]),
    ("gapx_server_session_submit", C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_uint, C.POINTER(C.c_int64)]),
    ("gapx_server_session_stats", C.c_int, [C.c_void_p, C.c_int64] + [C.POINTER(C.c_int64)] * 3),
    ("gapx_server_session_close", C.c_int, [C.c_void_p, C.c_int64]),
    ("gapx_connect", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double,
                               C.POINTER(C.c_void_p)]),
    ("gapx_nl_counts", C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p]),
    ("gapx_nl_export", C.c_int, [C.c_void_p] + [C.c_void_p] * 8),
    ("gapx_nl_free", C.c_int, [C.c_void_p]),
    ("gapx_descriptors", C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]),
    ("gapx_model_set_autotune", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_model_set_fused_schedule", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_model_set_graphs", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_model_set_nl_reuse", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_last_nl_rebuilt", C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    ("gapx_model_tuning", C.c_int, [C.c_void_p, C.c_int, C.c_int] + [C.c_void_p] * 6 + [C.POINTER(C.c_int)]),
    ("gapx_model_tuning_split", C.c_int, [C.c_void_p, C.c_int, C.c_int] + [C.c_void_p] * 7 + [C.POINTER(C.c_int)]),
    ("gapx_model_set_timing", C.c_int, [C.c_void_p, C.c_int]),
    ("gapx_last_timings", C.c_int, [C.c_void_p, C.c_void_p]),
    ("gapx_timing_totals", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
]

EXPORTED = [s[0] for s in _SIGS]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GapxError("libgapx.so not built (%s): run __graft_entry__.build()" % LIB_PATH)
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise GapxError("cannot load %s: %s" % (LIB_PATH, e))
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    buf = C.create_string_buffer(2048)
    lib().gapx_last_error(buf, 2048)
    return buf.value.decode(errors="replace")


def check(status, what=""):
    if status != 0:
        raise GapxError("%s failed (status %d): %s" % (what, status, last_error()), status)


def ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return C.c_void_p(a)
    return a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
