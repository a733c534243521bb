"""ctypes wrapper of oracle/liboracle.so -- the CPU checker (test infrastructure)."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                           stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.oracle_backward_error.restype = ctypes.c_double
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def factor(bcsc, a2x, fp32_updates=False, nthreads=0):
    """L values (xsize) of the oracle factorization of A2 = tril(P A P^T)."""
    L = np.zeros(bcsc.xsize, dtype=np.float64)
    a2x = np.ascontiguousarray(a2x, dtype=np.float64)
    info = lib().oracle_factor(
        ctypes.c_int(bcsc.n), _p(bcsc.a2p), _p(bcsc.a2i), _p(a2x), _p(bcsc.p), _p(bcsc.s),
        _p(bcsc.i_ptr), _p(L), _p(bcsc.super), ctypes.c_int(bcsc.nsuper), _p(bcsc.prunePtr),
        _p(bcsc.pruneSet), ctypes.c_int(int(fp32_updates)), ctypes.c_int(nthreads))
    return L, info


BLAS_PATH = os.environ.get("SWC_ORACLE_BLAS", "/opt/conda/lib/libmkl_rt.so.1")


def blas_available():
    """True when the BLAS-3 oracle can load its BLAS/LAPACK library."""
    return os.path.exists(BLAS_PATH) and lib().oracle_blas_load(BLAS_PATH.encode()) == 0


def factor_blas(bcsc, a2x, nthreads=0):
    """The BLAS-3 form of the oracle (the reference's dsyrk/dgemm/dpotrf/dtrsm calls)."""
    if not blas_available():
        raise RuntimeError(f"no BLAS for the oracle at {BLAS_PATH}")
    L = np.zeros(bcsc.xsize, dtype=np.float64)
    a2x = np.ascontiguousarray(a2x, dtype=np.float64)
    info = lib().oracle_factor_blas(
        ctypes.c_int(bcsc.n), _p(bcsc.a2p), _p(bcsc.a2i), _p(a2x), _p(bcsc.p), _p(bcsc.s),
        _p(bcsc.i_ptr), _p(L), _p(bcsc.super), ctypes.c_int(bcsc.nsuper), _p(bcsc.prunePtr),
        _p(bcsc.pruneSet), ctypes.c_int(nthreads))
    if info < 0:
        raise RuntimeError(f"oracle_factor_blas failed ({info})")
    return L, info


def lsolve(bcsc, L, x):
    x = np.asfortranarray(np.array(x, dtype=np.float64).reshape(bcsc.n, -1))
    lib().oracle_lsolve(ctypes.c_int(bcsc.n), _p(bcsc.p), _p(bcsc.s), _p(bcsc.i_ptr), _p(L),
                        _p(bcsc.super), ctypes.c_int(bcsc.nsuper), _p(x),
                        ctypes.c_int(x.shape[1]), ctypes.c_int(bcsc.n))
    return x


def ltsolve(bcsc, L, x):
    x = np.asfortranarray(np.array(x, dtype=np.float64).reshape(bcsc.n, -1))
    lib().oracle_ltsolve(ctypes.c_int(bcsc.n), _p(bcsc.p), _p(bcsc.s), _p(bcsc.i_ptr), _p(L),
                         _p(bcsc.super), ctypes.c_int(bcsc.nsuper), _p(x),
                         ctypes.c_int(x.shape[1]), ctypes.c_int(bcsc.n))
    return x


def solve(bcsc, L, b):
    """x = A^{-1} b in the original ordering (perm, L, L^T, perm^T)."""
    b = np.asarray(b, dtype=np.float64)
    vec = b.ndim == 1
    B = b.reshape(bcsc.n, -1)
    y = B[bcsc.Perm, :]
    y = ltsolve(bcsc, L, lsolve(bcsc, L, y))
    x = np.empty_like(y)
    x[bcsc.Perm, :] = y
    return x[:, 0] if vec else x


def backward_error(rp, ci, v, x, b):
    n = len(rp) - 1
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return lib().oracle_backward_error(ctypes.c_int(n), _p(rp), _p(ci), _p(v), _p(x), _p(b))


def to_sparse_L(bcsc, L):
    """Lower factor as a scipy CSC matrix (explicit zeros of the BCSC blocks kept)."""
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for s in range(bcsc.nsuper):
        c0, c1 = int(bcsc.super[s]), int(bcsc.super[s + 1])
        R = bcsc.s[bcsc.pi[s]:bcsc.pi[s + 1]]
        blk = L[bcsc.p[c0]:bcsc.p[c0] + len(R) * (c1 - c0)].reshape(c1 - c0, len(R))
        for j in range(c1 - c0):
            rows.append(R[j:])
            cols.append(np.full(len(R) - j, c0 + j))
            vals.append(blk[j, j:])
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(bcsc.n, bcsc.n))


def to_dense_L(bcsc, L):
    """Dense lower factor (n x n) from the BCSC values (small cases only)."""
    n = bcsc.n
    D = np.zeros((n, n))
    for s in range(bcsc.nsuper):
        c0, c1 = bcsc.super[s], bcsc.super[s + 1]
        rows = bcsc.s[bcsc.pi[s]:bcsc.pi[s + 1]]
        R = len(rows)
        blk = L[bcsc.p[c0]:bcsc.p[c0] + R * (c1 - c0)].reshape(c1 - c0, R).T
        D[np.ix_(rows, np.arange(c0, c1))] = blk
    return np.tril(D)
