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
    if info and prob.name.startswith("gmrf"):  # diagnostic for an SPD operator
        import os
        g = np.load(os.path.join(os.path.dirname(__file__), "golden", prob.name + ".npz"))
        names = {"perm": "Perm", "super": "super", "s": "s", "pi": "pi", "p": "p",
                 "i_ptr": "i_ptr", "pruneptr": "prunePtr", "pruneset": "pruneSet",
                 "a2_p": "a2p", "a2_i": "a2i"}
        diff = [k for k, a in names.items() if not np.array_equal(g[k], getattr(b, a))]
        print(f"oracle info {info} on {prob.name}; arrays differing from golden: {diff}; "
              f"a2x equal: {np.array_equal(g['a2_x'], a2x)}; "
              f"values sum {prob.values.sum():.17g}", flush=True)
        b2 = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
        print("fresh analysis equal:", all(np.array_equal(getattr(b, a), getattr(b2, a))
                                           for a in names.values()), flush=True)
        Lg, ig = O.factor(b, g["a2_x"], nthreads=1)
        print("oracle on golden values:", ig, flush=True)
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
