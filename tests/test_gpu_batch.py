"""GPU: a batch of uneven instances factored and solved in one launch sequence."""
import numpy as np
import pytest

import swcholesky_amd as S
import oracle_py as O
from swcholesky_amd.batch import Batch
from helpers import full_csr, relF
from test_batch import _instances

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_factor_matches_instances(mode):
    probs = _instances(12, seed=7, lo=16, hi=48)
    bt = Batch(probs)
    f = bt.factor
    f.set_precision(mode)
    assert f.factor(bt.values(probs)) == 0
    Lb = f.L_values()
    for k, p in enumerate(probs):
        b = S.analyze_p2(p.n, p.colptr, p.rowidx, p.perm)
        fi = S.Factor(b)
        fi.set_precision(mode)
        assert fi.factor(p.values) == 0
        Lk, _ = bt.instance_factor(k, Lb)
        # same supernodes, tasks and summation order: bitwise the instance's own factor
        assert np.array_equal(Lk, fi.L_values())
        if mode == 0:
            Lo, info = O.factor(b, b.a2_values(p.values))
            assert info == 0 and relF(Lk, Lo) < 1e-12


def test_batch_solve_and_refine():
    probs = _instances(10, seed=11, lo=16, hi=40)
    bt = Batch(probs)
    f = bt.factor
    rng = np.random.default_rng(0)
    rhs = rng.uniform(-1, 1, bt.n)
    for mode in (0, 1):
        f.set_precision(mode)
        assert f.factor(bt.values(probs)) == 0
        x, eta, iters = f.refine(rhs, tol=1e-12)
        assert eta[0] <= 1e-12
        # eta is the backward error of the block-diagonal system, whose norms are
        # the maxima over instances: every instance's residual meets that bound
        mats = [full_csr(p)[0] for p in probs]
        norm_a = max(abs(A).sum(axis=1).max() for A in mats)
        den = norm_a * np.abs(x).max() + np.abs(rhs).max()
        worst = max(np.abs(bk - A @ xk).max() for A, xk, bk in zip(mats, bt.split(x), bt.split(rhs)))
        assert worst / den <= 1e-12
        # and the oracle's own backward error of the whole batch agrees
        A, rp, ci, v = full_csr(_as_problem(bt, probs))
        assert O.backward_error(rp, ci, v, x, rhs) <= 1e-12


def _as_problem(bt, probs):
    """The batch as one Problem (block-diagonal lower triangle)."""
    from swcholesky_amd.problems import Problem
    return Problem(bt.n, bt.colptr, bt.rowidx, bt.values(probs), bt.perm, "batch", None)
