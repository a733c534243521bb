"""swcholesky_amd: MI355X-native numeric phase for swCholesky's supernodal Cholesky.

Host analysis (a restatement of the reference's analyze_p2 for a given ordering),
hand-written HIP kernels for gfx950 for the numeric factorization and the
triangular solves, all behind the C ABI of include/swcholesky_amd.h.
"""
from .api import BCSC, Factor, analyze_p2, blockedLsolve, sw_cholesky_left_par_05  # noqa: F401
from ._lib import SwcError, lib  # noqa: F401

__all__ = ["BCSC", "Factor", "analyze_p2", "blockedLsolve", "sw_cholesky_left_par_05", "SwcError",
           "lib"]
