"""Shared helpers of the parity tests."""
import numpy as np

import swcholesky_amd as S
from swcholesky_amd import problems as P
import oracle_py as O


def setup(prob):
    """(bcsc, a2x, oracle L, info) for a Problem."""
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    a2x = b.a2_values(prob.values)
    L, info = O.factor(b, a2x)
    return b, a2x, L, info


def relF(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def full_csr(prob):
    A = prob.scipy_full()
    A.sort_indices()
    return A, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


SMALL = [
    ("lap2d_const_100", lambda: P.laplacian2d_const(100)),
    ("lap2d_var_37", lambda: P.laplacian2d_var(37, 3)),
    ("lap3d_var_12", lambda: P.laplacian3d_var(12)),
    ("lap3d_var_20", lambda: P.laplacian3d_var(20)),
    ("vec3d_var_6", lambda: P.vector3d_var(6)),
    ("gmrf2d_32", lambda: P.gmrf2d(32)),
    ("dense_1100", lambda: P.dense_spd(1100)),
]
