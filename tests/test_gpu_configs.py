"""GPU parity at the full sizes of the stated workloads.

* 3D 7-point variable-coefficient Laplacian on 64^3 (n = 262,144): FP64 L against
  the oracle's L (relative Frobenius <= 1e-12), FP64 solve and both mixed modes
  refined to backward error <= 1e-12.
* 3-dof vector operator on 48^3 (n = 331,776): FP64 L against the oracle, then
  FP64 and mixed with 1, 8 and 32 right-hand sides ~ U(-1, 1), every column to
  eta <= 1e-12 by the oracle's checker.
* 4096 independent 2D instances, side ~ U{16..128}: FP64 and mixed, every
  instance to eta <= 1e-12 (its own norms), 32 sampled instances' L against the
  oracle.

The oracle here is its BLAS-3 form (the reference's dsyrk/dgemm/dpotrf/dtrsm
calls), pinned to the scalar oracle in tests/test_oracle.py.
"""
import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
from swcholesky_amd.batch import Batch
import oracle_py as O
from helpers import full_csr, relF

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not O.blas_available(), reason="no BLAS library for the oracle")]


class _Case:
    def __init__(self, prob, nthreads=0):
        self.prob = prob
        self.b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
        self.a2x = self.b.a2_values(prob.values)
        self.A, self.rp, self.ci, self.v = full_csr(prob)
        self.f = S.Factor(self.b)
        self.Lo, info = O.factor_blas(self.b, self.a2x, nthreads=nthreads)
        assert info == 0


def _eta(c, X, B):
    X = X.reshape(c.prob.n, -1)
    B = B.reshape(c.prob.n, -1)
    return np.array([O.backward_error(c.rp, c.ci, c.v, np.ascontiguousarray(X[:, j]),
                                      np.ascontiguousarray(B[:, j])) for j in range(X.shape[1])])


@pytest.fixture(scope="module")
def lap64():
    print("\n[configs] lap3d_var_64: analysis, upload, oracle factor", flush=True)
    return _Case(P.laplacian3d_var(64))


@pytest.mark.timeout(900)
def test_lap3d64_fp64_matches_oracle(lap64):
    c = lap64
    c.f.set_precision(0)
    assert c.f.factor(c.prob.values) == 0
    d = relF(c.f.L_values(), c.Lo)
    print(f"[configs] lap3d_var_64 FP64 relF(L) = {d:.3e}", flush=True)
    assert d <= 1e-12
    xs = np.random.default_rng(0).uniform(-1, 1, c.prob.n)
    rhs = c.A @ xs
    x = c.f.solve(rhs)
    eta = _eta(c, x, rhs)[0]
    print(f"[configs] lap3d_var_64 FP64 solve eta = {eta:.3e}", flush=True)
    assert eta <= 1e-12


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", [1, 2])
def test_lap3d64_mixed_refined(lap64, mode):
    c = lap64
    c.f.set_precision(mode)
    assert c.f.factor(c.prob.values) == 0
    xs = np.random.default_rng(0).uniform(-1, 1, c.prob.n)
    rhs = c.A @ xs
    x, eta, iters = c.f.refine(rhs, tol=1e-12)
    eo = _eta(c, x, rhs)[0]
    print(f"[configs] lap3d_var_64 mode {mode}: {iters} corrections, eta {eta[0]:.3e} "
          f"(oracle {eo:.3e})", flush=True)
    assert eta[0] <= 1e-12 and eo <= 1e-12
    c.f.set_precision(0)


@pytest.fixture(scope="module")
def vec48():
    print("\n[configs] vec3d_var_48: analysis, upload, oracle factor", flush=True)
    return _Case(P.vector3d_var(48))


@pytest.mark.timeout(900)
def test_vec48_fp64_matches_oracle(vec48):
    c = vec48
    c.f.set_precision(0)
    assert c.f.factor(c.prob.values) == 0
    d = relF(c.f.L_values(), c.Lo)
    print(f"[configs] vec3d_var_48 FP64 relF(L) = {d:.3e}", flush=True)
    assert d <= 1e-12


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("nrhs", [1, 8, 32])
def test_vec48_refine_every_column(vec48, mode, nrhs):
    c = vec48
    c.f.set_precision(mode)
    assert c.f.factor(c.prob.values) == 0
    B = np.random.default_rng(100 + nrhs).uniform(-1, 1, (c.prob.n, nrhs))
    X, eta, iters = c.f.refine(B, tol=1e-12)
    eo = _eta(c, X, B)
    print(f"[configs] vec3d_var_48 mode {mode} nrhs {nrhs}: {iters} corrections, "
          f"max eta {eta.max():.3e} (oracle {eo.max():.3e})", flush=True)
    assert np.all(eta <= 1e-12) and np.all(eo <= 1e-12)
    c.f.set_precision(0)


@pytest.fixture(scope="module")
def batch4096():
    print("\n[configs] batch of 4096 2D instances: generation and analysis", flush=True)
    rng = np.random.default_rng(3)
    sides = rng.integers(16, 129, size=4096)
    probs = [P.laplacian2d_var(int(sd), seed=None, rng=rng, wlo=1e-1, whi=1e1, shift=1e-4)
             for sd in sides]
    bt = Batch(probs)
    return bt, probs, bt.values(probs)


@pytest.mark.timeout(900)
@pytest.mark.parametrize("mode", [0, 1])
def test_batch4096_every_instance(batch4096, mode):
    bt, probs, vals = batch4096
    f = bt.factor
    f.set_precision(mode)
    assert f.factor(vals) == 0
    rhs = np.random.default_rng(0).uniform(-1, 1, bt.n)
    # refined a decade below the target, so that every instance's own backward
    # error (with its own norms, never larger than the batch's) meets 1e-12
    x, eta, iters = f.refine(rhs, tol=1e-13)
    worst = 0.0
    for p, xk, bk in zip(probs, bt.split(x), bt.split(rhs)):
        rp, ci, v = P.full_csr_from_lower(p.n, p.colptr, p.rowidx, p.values)
        worst = max(worst, O.backward_error(rp, ci, v, np.ascontiguousarray(xk),
                                            np.ascontiguousarray(bk)))
    print(f"[configs] batch4096 mode {mode}: {iters} corrections, worst instance eta "
          f"{worst:.3e}", flush=True)
    assert worst <= 1e-12
    if mode == 0:
        L = f.L_values()
        sample = np.random.default_rng(9).choice(len(probs), 32, replace=False)
        dmax = 0.0
        for k in sample:
            p = probs[k]
            b = S.analyze_p2(p.n, p.colptr, p.rowidx, p.perm)
            Lo, info = O.factor(b, b.a2_values(p.values))
            assert info == 0
            Lk, _ = bt.instance_factor(int(k), L)
            dmax = max(dmax, relF(Lk, Lo))
        print(f"[configs] batch4096: 32 sampled instances, max relF(L) {dmax:.3e}", flush=True)
        assert dmax <= 1e-12
    f.set_precision(0)
