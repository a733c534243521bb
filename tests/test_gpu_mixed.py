"""GPU: FP32-update factorization and FP64 refinement to 1e-12 backward error."""
import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
import oracle_py as O
from helpers import full_csr, relF, setup

pytestmark = pytest.mark.gpu

CASES = [("lap3d_var_16", lambda: P.laplacian3d_var(16)),
         ("vec3d_var_8", lambda: P.vector3d_var(8)),
         ("lap2d_const_100", lambda: P.laplacian2d_const(100))]


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_mixed_factor_is_a_perturbed_fp64_factor(name, make, mode):
    prob = make()
    b, a2x, L64, _ = setup(prob)
    f = S.Factor(b)
    f.set_precision(mode)
    assert f.factor(prob.values) == 0
    Lm = f.L_values()
    d = relF(Lm, L64)
    assert 1e-13 < d < 1e-3, d  # FP32 rounding is visible but small
    L32, info = O.factor(b, a2x, fp32_updates=True)
    assert relF(Lm, L32) < 10 * max(relF(L32, L64), 1e-7)


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("nrhs", [1, 8])
@pytest.mark.parametrize("name,make", CASES[:2], ids=[c[0] for c in CASES[:2]])
def test_mixed_refine_reaches_1e12(name, make, nrhs, method):
    prob = make()
    b, *_ = setup(prob)
    f = S.Factor(b)
    f.set_precision(1)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    B = np.random.default_rng(3).uniform(-1, 1, (prob.n, nrhs))
    X, eta, iters = f.refine(B, tol=1e-12, method=method)
    assert np.all(eta <= 1e-12) and iters >= 1
    for j in range(nrhs):
        assert O.backward_error(rp, ci, v, X[:, j], B[:, j]) <= 1.5e-12


def test_fp64_refine_and_gpu_backward_error():
    prob = P.laplacian3d_var(16)
    b, *_ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    rhs = A @ np.random.default_rng(0).uniform(-1, 1, prob.n)
    x, eta, iters = f.refine(rhs, tol=1e-12)
    assert iters == 0 and eta[0] <= 1e-14
    # at the rounding level both sides only agree on the order of magnitude
    assert f.backward_error(rhs, x)[0] <= 1e-14 and O.backward_error(rp, ci, v, x, rhs) <= 1e-14
    # a perturbed x has a residual far above rounding: both sides must agree closely
    xp = x + 1e-6 * np.random.default_rng(1).standard_normal(prob.n)
    assert np.isclose(f.backward_error(rhs, xp)[0], O.backward_error(rp, ci, v, xp, rhs),
                      rtol=1e-10)


@pytest.mark.parametrize("nrhs", [1, 4, 6])
def test_backward_error_columns_match_oracle(nrhs):
    """The residual kernel takes the right-hand sides four at a time against
    the same row entries; each column's eta matches the oracle's."""
    prob = P.vector3d_var(6)
    b, *_ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    rng = np.random.default_rng(nrhs)
    X = rng.uniform(-1, 1, (prob.n, nrhs))
    B = A @ X + 1e-6 * rng.standard_normal((prob.n, nrhs))
    eta = f.backward_error(B, X)
    for j in range(nrhs):
        assert np.isclose(eta[j], O.backward_error(rp, ci, v, X[:, j], B[:, j]), rtol=1e-10)
