"""Pins the CPU oracle: dense algebra, closed forms, reference known-answer shapes."""
import os

import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
import oracle_py as O
from helpers import full_csr, relF, setup

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("make", [lambda: P.laplacian2d_const(12), lambda: P.laplacian3d_var(8),
                                  lambda: P.vector3d_var(5), lambda: P.gmrf2d(20)])
def test_factor_equals_dense_cholesky(make):
    prob = make()
    b, a2x, L, info = setup(prob)
    assert info == 0
    Ld = O.to_dense_L(b, L)
    Ap = prob.scipy_full().toarray()[np.ix_(b.Perm, b.Perm)]
    assert relF(Ld @ Ld.T, Ap) < 1e-14
    assert relF(Ld, np.linalg.cholesky(Ap)) < 1e-11


def test_forward_solve_closed_form():
    """b = L 1 => L y = b gives y = 1 (rhsInitBlocked / testTriangular, Util.h:277-306)."""
    prob = P.laplacian2d_const(40)
    b, a2x, L, _ = setup(prob)
    Ld = O.to_dense_L(b, L)
    rhs = Ld @ np.ones(b.n)
    y = O.lsolve(b, L, rhs)[:, 0]
    assert np.max(np.abs(y - 1)) < 1e-12
    z = O.ltsolve(b, L, Ld.T @ np.ones(b.n))[:, 0]
    assert np.max(np.abs(z - 1)) < 1e-12


def test_config1_golden_solution():
    g = np.load(os.path.join(GOLDEN, "c1_lap2d_const_100.npz"))
    prob = P.laplacian2d_const(100)
    b, a2x, L, _ = setup(prob)
    norms = np.array([np.linalg.norm(L[b.p[b.super[s]]:b.p[b.super[s]] + (b.pi[s + 1] - b.pi[s])
                                       * (b.super[s + 1] - b.super[s])]) for s in range(b.nsuper)])
    assert np.allclose(norms, g["oracle_snorm"], rtol=1e-13)
    x = O.solve(b, L, g["b"])
    assert np.allclose(x, g["oracle_x"], rtol=0, atol=1e-12)
    assert np.max(np.abs(x - g["xstar"])) < 1e-9
    A, rp, ci, v = full_csr(prob)
    assert O.backward_error(rp, ci, v, x, g["b"]) < 1e-15


def _dense_as_sparse(M):
    n = M.shape[0]
    Lw = np.tril(M)
    cols, rows = np.nonzero(Lw.T)
    return P.Problem(n, *P.lower_csc_from_coo(n, rows, cols, Lw[rows, cols]), np.arange(n),
                     "dense")


@pytest.mark.parametrize("n", [64, 343])
def test_single_supernode_potrf_known_answer(n):
    """test_dpotrf.c:13-33 shape/init: a = 0.2 + (i%100)*0.001, +20 on the diagonal."""
    a = 0.2 + (np.arange(n * n) % 100) * 0.001
    M = a.reshape(n, n)
    M = np.tril(M) + np.tril(M, -1).T + 20 * n * 0.01 * np.eye(n)
    M = (M + M.T) / 2 + 20 * np.eye(n)
    prob = _dense_as_sparse(M)
    b, a2x, L, info = setup(prob)
    assert b.nsuper == 1 and info == 0
    Ld = O.to_dense_L(b, L)
    Pm = M[np.ix_(b.Perm, b.Perm)]
    assert np.max(np.abs(Ld - np.linalg.cholesky(Pm))) < 1e-12


def test_not_spd_info():
    prob = P.laplacian2d_var(12, 1, shift=-3.0)
    b, a2x, L, info = setup(prob)
    assert info > 0


def test_fp32_updates_are_close():
    prob = P.laplacian2d_var(30, 2, shift=1e-2)
    b, a2x, L64, _ = setup(prob)
    L32, info = O.factor(b, a2x, fp32_updates=True)
    assert info == 0
    assert 1e-12 < relF(L32, L64) < 1e-4
