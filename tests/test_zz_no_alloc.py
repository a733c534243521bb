"""swc_solve_noalloc (the solve entry point that allocates nothing) through the C
ABI, against swc_solve on the same handle: bitwise-equal solutions (same
kernels, same order), on two calls in a row, the second with a smaller input,
and a clear error for a request beyond its fixed workspace."""
import ctypes

import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import _lib, problems as P
import oracle_py as O
from helpers import full_csr

pytestmark = pytest.mark.gpu


def _solve_noalloc(f, X):
    """In-place no-alloc solve of the device tensor X (nrhs, n)."""
    nrhs = 1 if X.dim() == 1 else X.shape[0]
    return _lib.lib().swc_solve_noalloc(f._h, ctypes.c_void_p(X.data_ptr()), nrhs, f.n)


def _case(prob):
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    return f


def test_noalloc_solve_matches_solve_twice_second_smaller():
    import torch
    big = P.laplacian3d_var(20)
    small = P.laplacian3d_var(12)
    fb, fs = _case(big), _case(small)
    rng = np.random.default_rng(11)
    for f, prob, nrhs in ((fb, big, 5), (fs, small, 2), (fb, big, 1)):
        print(f"[noalloc] n={prob.n} nrhs={nrhs}", flush=True)
        B = rng.uniform(-1, 1, (nrhs, prob.n))
        Xd = torch.from_numpy(B.copy()).cuda()
        assert _solve_noalloc(f, Xd) == 0, _lib.lib().swc_last_error()
        Xn = Xd.cpu().numpy()
        Xr = f.solve(B.T.copy()).reshape(prob.n, nrhs)
        assert np.array_equal(Xn.T, Xr)
        _, rp, ci, v = full_csr(prob)
        for j in range(nrhs):
            eta = O.backward_error(rp, ci, v, np.ascontiguousarray(Xn[j]), np.ascontiguousarray(B[j]))
            assert eta <= 1e-14


def test_noalloc_rejects_requests_beyond_its_workspace():
    import torch
    prob = P.laplacian3d_var(12)
    f = _case(prob)
    nrhs = 4000  # (2n + vsize + ...) * 8 bytes * 4000 is well past 128 MiB
    X = torch.zeros((nrhs, prob.n), dtype=torch.float64, device="cuda")
    rc = _solve_noalloc(f, X)
    assert rc != 0
    assert b"fixed workspace" in _lib.lib().swc_last_error()
    # the handle still solves afterwards
    x = np.ones(prob.n)
    Xd = torch.from_numpy(x.copy()).cuda()
    assert _solve_noalloc(f, Xd) == 0
    assert np.array_equal(Xd.cpu().numpy(), f.solve(x))
