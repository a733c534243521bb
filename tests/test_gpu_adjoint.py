"""GPU adjoint: selected inverse, logdet and gradients of logdet(Q) + b^T Q^{-1} b.

Pinned by dense linear algebra (numpy inverse, torch.linalg autograd) on small
grids, central finite differences (h = 1e-6) and solves at the full GMRF size."""
import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P

pytestmark = pytest.mark.gpu


def _factor(prob):
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    return b, f


# dense_spd(200) is one supernode of 200 columns and lap3d 12^3 has separators of up
# to 144 columns: T = L11^{-1} then spans several 64-row blocks (rows above a block
# applied from LDS), beside the narrow supernodes of the other cases
@pytest.mark.parametrize("make", [lambda: P.gmrf2d(32), lambda: P.laplacian3d_var(8),
                                  lambda: P.vector3d_var(4), lambda: P.laplacian2d_const(30),
                                  lambda: P.dense_spd(200, seed=7),
                                  lambda: P.laplacian3d_var(12)])
def test_selected_inverse_matches_dense(make):
    prob = make()
    b, f = _factor(prob)
    Z = f.selected_inverse()
    Ap = prob.scipy_full().toarray()[np.ix_(b.Perm, b.Perm)]
    Ainv = np.linalg.inv(Ap)
    err = 0.0
    for s in range(b.nsuper):
        c0, c1 = b.super[s], b.super[s + 1]
        rows = b.s[b.pi[s]:b.pi[s + 1]]
        R = len(rows)
        blk = Z[b.p[c0]:b.p[c0] + R * (c1 - c0)].reshape(c1 - c0, R).T
        ref = Ainv[np.ix_(rows, np.arange(c0, c1))]
        ref = np.where(rows[:, None] >= np.arange(c0, c1)[None, :], ref, 0.0)
        err = max(err, np.abs(blk - ref).max())
    # a computed inverse carries errors ~ cond(A) eps |A^-1| (numpy's as well)
    tol = max(1e-10, 20 * np.linalg.cond(Ap) * np.finfo(float).eps)
    assert err <= tol * np.abs(Ainv).max(), (err, tol)
    sign, ld = np.linalg.slogdet(Ap)
    assert sign > 0 and abs(f.logdet() - ld) <= 1e-10 * abs(ld)


def test_gradient_matches_dense_torch_linalg():
    import torch
    prob = P.gmrf2d(32)
    b, f = _factor(prob)
    rng = np.random.default_rng(0)
    rhs = rng.standard_normal(prob.n)
    x = f.solve(rhs)
    g = f.gradient(x, alpha=1.0, beta=1.0)
    # dense torch reference: A built from the stored values, autograd through linalg
    vals = torch.tensor(prob.values, dtype=torch.float64, requires_grad=True)
    cols = np.repeat(np.arange(prob.n), np.diff(prob.colptr))
    rows = prob.rowidx
    A = torch.zeros(prob.n, prob.n, dtype=torch.float64)
    A = A.index_put((torch.from_numpy(rows.astype(np.int64)), torch.from_numpy(cols)), vals)
    off = torch.from_numpy((rows != cols))
    At = A + (A * 0).index_put((torch.from_numpy(cols[rows != cols]),
                                torch.from_numpy(rows[rows != cols].astype(np.int64))),
                               vals[off])
    bt = torch.from_numpy(rhs)
    loss = torch.logdet(At) + bt @ torch.linalg.solve(At, bt)
    loss.backward()
    ref = vals.grad.numpy()
    assert np.abs(g - ref).max() <= 1e-9 * np.abs(ref).max()


def test_gradient_central_differences():
    prob = P.gmrf2d(32)
    b, f = _factor(prob)
    rhs = np.random.default_rng(1).standard_normal(prob.n)

    def loss(vals):
        assert f.factor(vals) == 0
        x = f.solve(rhs)
        return f.logdet() + rhs @ x

    assert f.factor(prob.values) == 0
    g = f.gradient(f.solve(rhs), 1.0, 1.0)
    h = 1e-6
    picks = np.random.default_rng(2).choice(prob.nnz, 12, replace=False)
    for q in picks:
        vp = prob.values.copy(); vp[q] += h
        vm = prob.values.copy(); vm[q] -= h
        fd = (loss(vp) - loss(vm)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(g[q])), (q, fd, g[q])


def test_autograd_function_gmrf_full_size():
    import torch
    from swcholesky_amd.autograd import logdet_quad
    prob = P.gmrf2d(256)  # the GMRF configuration, n = 65,536
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    f = S.Factor(b)
    vals = torch.tensor(prob.values, dtype=torch.float64, device="cuda", requires_grad=True)
    rhs = torch.tensor(np.random.default_rng(4).standard_normal(prob.n), device="cuda")
    out = logdet_quad(vals, rhs, f)
    out.backward()
    g = vals.grad.cpu().numpy()
    # selected-inverse entries against solves: Z_ij = e_i^T A^{-1} e_j
    Z = f.selected_inverse()
    rng = np.random.default_rng(5)
    for _ in range(6):
        s = int(rng.integers(b.nsuper))
        c = int(rng.integers(b.super[s], b.super[s + 1]))
        rows = b.s[b.pi[s]:b.pi[s + 1]]
        r = int(rows[rng.integers(c - b.super[s], len(rows))])
        e = np.zeros(prob.n); e[b.Perm[c]] = 1.0
        col = f.solve(e)
        z = Z[b.p[c] + rows.searchsorted(r)]
        assert abs(col[b.Perm[r]] - z) <= 1e-10 * abs(col[b.Perm[c]])
    # central differences of the loss on a few values at full size
    base = prob.values
    for q in np.random.default_rng(6).choice(prob.nnz, 3, replace=False):
        def loss(v):
            t = torch.tensor(v, device="cuda")
            return float(logdet_quad(t, rhs, f))
        vp = base.copy(); vp[q] += 1e-6
        vm = base.copy(); vm[q] -= 1e-6
        fd = (loss(vp) - loss(vm)) / 2e-6
        assert abs(fd - g[q]) <= 1e-5 * max(1.0, abs(g[q])), (q, fd, g[q])
