"""ctypes binding of libswcholesky_amd.so (the C ABI of include/swcholesky_amd.h).

The library is built in-tree by ``python -m swcholesky_amd.build``.  There is no
fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswcholesky_amd.so")

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); every symbol declared in include/swcholesky_amd.h
SIGNATURES = {
    "swc_last_error": (ctypes.c_char_p, []),
    "swc_analyze": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "swc_analyze_batch": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int,
                                  ctypes.POINTER(c_vp)]),
    "swc_symbolic_free": (None, [c_vp]),
    "swc_nd_order": (c_int, [c_int, c_vp, c_vp, c_vp]),
    "swc_symbolic_sizes": (c_int, [c_vp, c_vp, ctypes.POINTER(c_dbl)]),
    "swc_symbolic_copy": (c_int, [c_vp, c_int, c_vp]),
    "sw_cholesky_left_par_05": (c_int, [c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_int, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_int, c_vp,
                                        c_vp, c_int, c_int, c_int, c_int, c_vp]),
    "blockedLsolve": (c_int, [c_int, c_vp, c_vp, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_vp]),
    "swc_create": (c_int, [ctypes.POINTER(c_vp)]),
    "swc_destroy": (c_int, [c_vp]),
    "swc_upload": (c_int, [c_vp, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                           c_vp, c_i64]),
    "swc_upload_symbolic": (c_int, [c_vp, c_vp]),
    "swc_set_perm": (c_int, [c_vp, c_vp]),
    "swc_factor": (c_int, [c_vp, c_vp, c_int, c_ip]),
    "swc_get_factor": (c_int, [c_vp, c_vp, c_int]),
    "swc_factor_nnz": (c_int, [c_vp, ctypes.POINTER(c_i64)]),
    "swc_factor_csc": (c_int, [c_vp, c_vp, c_vp, c_vp, c_int]),
    "bcsc2csc": (c_int, [ctypes.c_size_t, ctypes.c_size_t, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                         c_vp]),
    "swc_solve": (c_int, [c_vp, c_vp, c_int, c_i64, c_int]),
    "swc_solve_noalloc": (c_int, [c_vp, c_vp, c_int, c_i64]),
    "swc_plan_stats": (c_int, [c_vp, c_vp, ctypes.POINTER(c_dbl)]),
    "swc_level_timings": (c_int, [c_vp, c_vp, c_int, c_vp]),
    "swc_set_profiling": (c_int, [c_vp, c_int]),
    "swc_set_graphs": (c_int, [c_vp, c_int]),
    "swc_selected_inverse": (c_int, [c_vp, c_vp, c_int]),
    "swc_logdet": (c_int, [c_vp, ctypes.POINTER(c_dbl)]),
    "swc_gradient": (c_int, [c_vp, c_vp, c_int, c_dbl, c_dbl, c_vp, c_int]),
    "swc_set_precision": (c_int, [c_vp, c_int]),
    "swc_set_plan_params": (c_int, [c_vp, c_vp, c_int]),
    "swc_set_schedule": (c_int, [c_vp, c_int, c_vp, c_vp, c_vp]),
    "swc_refine": (c_int, [c_vp, c_vp, c_vp, c_int, c_i64, c_int, c_dbl, c_int, c_int, c_vp, c_ip]),
    "swc_residual": (c_int, [c_vp, c_vp, c_vp, c_int, c_i64, c_int, c_vp]),
    "swc_get_timings": (c_int, [c_vp, c_vp]),
    "swc_plan_build": (c_int, [c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                               ctypes.POINTER(c_vp)]),
    "swc_plan_build_params": (c_int, [c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int,
                                      c_int, c_vp, c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "swc_plan_free": (None, [c_vp]),
    "swc_plan_sizes": (c_int, [c_vp, c_vp, ctypes.POINTER(c_dbl)]),
    "swc_plan_export": (c_int, [c_vp, c_int, c_vp]),
    "swc_plan_struct_sizes": (c_int, [c_vp]),
}

_lib = None


class SwcError(RuntimeError):
    """A C ABI call returned a nonzero status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load (once) and return the shared library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m swcholesky_amd.build`")
        # torch ships its own libamdhip64.so.7; loading it first makes this library
        # bind to the same HIP runtime, so device pointers are shared safely.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    if rc != 0:
        raise SwcError(rc, lib().swc_last_error().decode(errors="replace"))
    return rc


def ptr(a):
    """Data pointer of a numpy array (or None)."""
    return None if a is None else ctypes.c_void_p(a.ctypes.data)
