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


# ---- backward error: the pass criterion of every GPU solve test ----------

def test_backward_error_hand_cases():
    """eta = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) on 2 x 2 cases worked
    out by hand: A = [[4, 1], [1, 3]], x = (1, -2): A x = (2, -5), ||A||_inf = 5,
    ||x||_inf = 2.  With b = 0: r = (-2, 5), eta = 5 / 10.  With b = (1, 1):
    r = (-1, 6), eta = 6 / 11 (a 1-norm, 2-norm or missing ||b|| term gives another
    value in each case)."""
    rp = np.array([0, 2, 4], np.int32)
    ci = np.array([0, 1, 0, 1], np.int32)
    v = np.array([4.0, 1.0, 1.0, 3.0])
    x = np.array([1.0, -2.0])
    assert O.backward_error(rp, ci, v, x, np.zeros(2)) == 0.5
    assert O.backward_error(rp, ci, v, x, np.ones(2)) == 6.0 / 11.0
    # the exact solution has eta 0
    assert O.backward_error(rp, ci, v, x, np.array([2.0, -5.0])) == 0.0


@pytest.mark.parametrize("make", [lambda: P.laplacian3d_var(10), lambda: P.vector3d_var(6),
                                  lambda: P.gmrf2d(24)])
def test_backward_error_matches_numpy(make):
    """The oracle's eta against an independent scipy/numpy evaluation of the same
    definition, on perturbed solutions (residuals far above rounding)."""
    prob = make()
    A, rp, ci, v = full_csr(prob)
    rng = np.random.default_rng(5)
    for scale in (1e-3, 1e-8, 1.0):
        x = rng.uniform(-1, 1, prob.n)
        b = A @ x + scale * rng.standard_normal(prob.n)
        xp = x + scale * rng.standard_normal(prob.n)
        r = b - A @ xp
        ref = np.abs(r).max() / (abs(A).sum(axis=1).max() * np.abs(xp).max() + np.abs(b).max())
        assert np.isclose(O.backward_error(rp, ci, v, xp, b), ref, rtol=1e-13, atol=0)


# ---- FP32-update mode: one update against a float32 product --------------

def _arrow(vals):
    """Lower triangle of a symmetric arrowhead [[d0, 0, b0], [0, d1, b1], [b0, b1, d2]]
    in the identity ordering.  Without unconditional merges of small supernodes
    (nrelax = 0, zrelax = 0), merging columns 1 and 2 adds no zero and merging the
    leaf would, so the partition is {0}, {1, 2}: one update, from the one-column leaf {0} (rows 0, 2)
    onto entry (2, 2) of the second supernode."""
    d0, d1, d2, b0, b1 = vals
    n = 3
    cp = np.array([0, 2, 4, 5], np.int32)
    ri = np.array([0, 2, 1, 2, 2], np.int32)
    va = np.array([d0, b0, d1, b1, d2])
    return P.Problem(n, cp, ri, va, np.arange(n), "arrow")


def test_fp32_update_against_float32_product():
    """In the FP32-update mode the update of entry (2, 2) is fl32(f f), f = fl32(l20)
    the float32-rounded leaf entry, subtracted in FP64; in FP64 it is l20^2 in FP64.
    The target's own POTRF is FP64 in both modes:  L22 = sqrt((d2 - c) - l21^2),
    l21 = b1 / sqrt(d1)."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        d0, d1 = rng.uniform(2, 5, 2)
        b0, b1 = rng.uniform(-1, 1, 2)
        d2 = 3.0 + rng.uniform()
        prob = _arrow((d0, d1, d2, b0, b1))
        b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm, nrelax=(0, 0, 0),
                         zrelax=(0.0, 0.0, 0.0))
        assert list(b.super) == [0, 1, 3] and list(b.Perm) == [0, 1, 2]
        assert list(b.pruneSet[b.prunePtr[1]:b.prunePtr[2]]) == [0]
        a2x = b.a2_values(prob.values)
        L64, i64 = O.factor(b, a2x)
        L32, i32 = O.factor(b, a2x, fp32_updates=True)
        assert i64 == 0 and i32 == 0
        # the leaf has no update: identical in both modes, and l20 = b0 / sqrt(d0) to
        # rounding (the TRSM multiplies by the reciprocal of the pivot)
        l20 = L64[b.p[0] + 1]
        assert L32[b.p[0] + 1] == l20 and np.isclose(l20, b0 / np.sqrt(d0), rtol=1e-15)
        l21 = b1 / np.sqrt(d1)
        root = b.p[2] + 1  # column 2 = second column of the target, its diagonal row
        f = np.float32(l20)
        c32 = float(np.float32(float(f) * float(f)))  # exact product, rounded once
        assert L32[root] == np.sqrt((d2 - c32) - l21 * l21)
        assert L64[root] == np.sqrt((d2 - l20 * l20) - l21 * l21)
        assert L32[root] != L64[root]


# ---- BLAS-3 form of the oracle -------------------------------------------

needs_blas = pytest.mark.skipif(not O.blas_available(), reason="no BLAS library for the oracle")


@needs_blas
@pytest.mark.parametrize("make", [lambda: P.laplacian2d_const(40), lambda: P.laplacian3d_var(14),
                                  lambda: P.vector3d_var(7), lambda: P.dense_spd(300),
                                  lambda: P.gmrf2d(32)])
def test_blas_oracle_matches_scalar_oracle(make):
    prob = make()
    b, a2x, L, info = setup(prob)
    Lb, ib = O.factor_blas(b, a2x, nthreads=4)
    assert info == 0 and ib == 0
    assert relF(Lb, L) < 1e-13


@needs_blas
def test_blas_oracle_residual_and_info():
    """||P A P^T v - L (L^T v)|| / (||A|| ||v||) at rounding level for random v (a
    dropped update or a wrong index shows up far above it), and the same first
    failing column as the scalar oracle on an indefinite operator."""
    prob = P.laplacian3d_var(24)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    Lb, ib = O.factor_blas(b, b.a2_values(prob.values))
    assert ib == 0
    Ap = prob.scipy_full()[b.Perm][:, b.Perm]
    Ls = O.to_sparse_L(b, Lb)
    v = np.random.default_rng(2).standard_normal(b.n)
    err = np.linalg.norm(Ap @ v - Ls @ (Ls.T @ v)) / (abs(Ap).sum(axis=1).max() * np.linalg.norm(v))
    assert err < 1e-15
    bad = P.laplacian2d_var(20, 5, shift=-2.0)
    bb, a2x, _, info = setup(bad)
    assert info > 0 and O.factor_blas(bb, a2x)[1] == info
