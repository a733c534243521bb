"""The C-ABI library loads and exports every symbol include/swcholesky_amd.h declares."""
import ctypes
import os
import re

import swcholesky_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "swcholesky_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|void|char|double)\s*\*?\s*(\w+)\s*\(", text, re.M)
    return sorted(set(names))


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(S._lib.LIB_PATH)
    names = declared_functions()
    assert "sw_cholesky_left_par_05" in names and "swc_factor" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_functions()) <= set(S._lib.SIGNATURES)


def test_error_message_roundtrip():
    import numpy as np
    import pytest
    with pytest.raises(S.SwcError) as e:
        S.analyze_p2(3, np.array([0, 1, 2, 3]), np.array([0, 1, 2]), np.array([0, 0, 1]))
    assert "perm" in str(e.value)


def test_blockedLsolve_null_inputs_return_zero():
    """The reference returns 0 when Lp, Li or x is null (Triangular_BCSC.h:23)."""
    lib = S.lib()
    assert lib.blockedLsolve(1, None, None, None, 0, None, None, None, 1, None) == 0
