"""Generators, orderings and the reference's file formats."""
import numpy as np
import scipy.sparse.linalg as sla

from swcholesky_amd import problems as P


def test_generators_spd_and_sorted():
    for prob in [P.laplacian2d_const(10), P.laplacian2d_var(9, 3), P.laplacian3d_var(5),
                 P.vector3d_var(3), P.gmrf2d(8)]:
        A = prob.scipy_full()
        assert (A - A.T).count_nonzero() == 0
        ev = np.linalg.eigvalsh(A.toarray())
        assert ev.min() > 0
        cp, ri = prob.colptr, prob.rowidx
        for j in range(prob.n):
            col = ri[cp[j]:cp[j + 1]]
            assert col[0] == j and np.all(np.diff(col) > 0)
        assert sorted(prob.perm) == list(range(prob.n))


def test_config_sizes():
    assert P.laplacian2d_const(100).nnz == 29800
    c2 = P.laplacian3d_var(16)
    assert c2.n == 16 ** 3 and c2.nnz == 16 ** 3 + 3 * 15 * 16 * 16
    v = P.vector3d_var(4)
    assert v.n == 3 * 64


def test_nd_order_separators_last():
    perm = P.grid_nd_order((7, 7))
    assert perm[-1] in {21, 22, 23, 24, 25, 26, 27}  # middle row of a 7x7 grid last


def test_mtx_and_perm_roundtrip(tmp_path):
    prob = P.laplacian2d_var(6, 3)
    f = tmp_path / "a.mtx"
    P.write_mtx(str(f), prob.n, prob.colptr, prob.rowidx, prob.values)
    n, cp, ri, va = P.read_mtx(str(f))
    assert n == prob.n and np.array_equal(cp, prob.colptr) and np.array_equal(ri, prob.rowidx)
    assert np.array_equal(va, prob.values)
    pf = tmp_path / "a.perm"
    P.write_perm(str(pf), prob.perm)
    assert np.array_equal(P.read_perm(str(pf), prob.n), prob.perm)
    assert pf.stat().st_size == 8 * prob.n


def test_gmrf_precision_definition():
    prob = P.gmrf2d(6, tau=2.0, kappa2=0.5)
    A = prob.scipy_full()
    lap = A / 2.0 - 0.5 * np.eye(prob.n)
    assert np.allclose(np.asarray(lap.sum(axis=1)).ravel(), 0)


def test_bcsc2csc_export_layout():
    """bcsc2csc (the reference's export, common/Util.h:311-338) against a
    column-by-column numpy restatement: column j holds its supernode's rows from
    its own diagonal down with their stored values; L times its transpose
    reproduces P A P^T."""
    import numpy as np
    import swcholesky_amd as S
    from swcholesky_amd import api
    import oracle_py as O
    prob = P.laplacian3d_var(10)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    L, info = O.factor(b, b.a2_values(prob.values))
    assert info == 0
    Cp, Ci, Cx = api.bcsc2csc(b, L)
    exp_p, exp_i, exp_x = [0], [], []
    for s in range(b.nsuper):
        c0, c1 = int(b.super[s]), int(b.super[s + 1])
        R = b.s[b.pi[s]:b.pi[s + 1]]
        for j in range(c0, c1):
            d = j - c0
            exp_i.extend(R[d:])
            exp_x.extend(L[b.p[j] + d:b.p[j + 1]])
            exp_p.append(len(exp_i))
    assert np.array_equal(Cp, exp_p) and np.array_equal(Ci, exp_i) and np.array_equal(Cx, exp_x)
    import scipy.sparse as sp
    Lc = sp.csc_matrix((Cx, Ci, Cp), shape=(b.n, b.n))
    A = prob.scipy_full()[b.Perm][:, b.Perm]
    assert abs(Lc @ Lc.T - A).max() < 1e-10 * abs(A).max()


def test_problem_round_trips_through_mtx_and_perm_files(tmp_path):
    """A problem written as MatrixMarket lower triangle + binary size_t perm (the
    files the reference reads, Util.h:77-223, sw_choleskyTest01.cpp:79) reads
    back to the same arrays and the same analysis."""
    import numpy as np
    import swcholesky_amd as S
    prob = P.laplacian3d_var(8)
    P.write_mtx(str(tmp_path / "a.mtx"), prob.n, prob.colptr, prob.rowidx, prob.values)
    P.write_perm(str(tmp_path / "a.perm"), prob.perm)
    n, cp, ri, v = P.read_mtx(str(tmp_path / "a.mtx"))
    perm = P.read_perm(str(tmp_path / "a.perm"), n)
    assert n == prob.n and np.array_equal(cp, prob.colptr) and np.array_equal(ri, prob.rowidx)
    assert np.array_equal(v, prob.values) and np.array_equal(perm, prob.perm)
    b1 = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    b2 = S.analyze_p2(n, cp, ri, perm)
    assert np.array_equal(b1.s, b2.s) and np.array_equal(b1.pruneSet, b2.pruneSet)


def test_nd_order_of_general_patterns():
    """swc_nd_order (the METIS_NodeND stand-in, LSparsity.h:597): a bijection, and
    a fill-reducing one -- on grids no more fill than the geometric
    nested dissection, on an unstructured Delaunay mesh a fraction of reverse
    Cuthill-McKee's fill and of the natural order's."""
    import numpy as np
    from scipy.sparse.csgraph import reverse_cuthill_mckee
    import swcholesky_amd as S

    def nnz_l(prob, perm):
        b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, perm)
        return int(np.sum(b.ColCount))

    for prob in (P.laplacian2d_var(64, 3), P.laplacian3d_var(16), P.vector3d_var(6)):
        perm = P.nd_order(prob.n, prob.colptr, prob.rowidx)
        assert np.array_equal(np.sort(perm), np.arange(prob.n))
        assert nnz_l(prob, perm) <= 1.05 * nnz_l(prob, prob.perm)
    m = P.mesh2d_random(6000)
    assert np.array_equal(np.sort(m.perm), np.arange(m.n))
    rcm = reverse_cuthill_mckee(m.scipy_full(), symmetric_mode=True).astype(np.int64)
    assert nnz_l(m, m.perm) < 0.5 * nnz_l(m, rcm)
    assert nnz_l(m, m.perm) < 0.1 * nnz_l(m, np.arange(m.n, dtype=np.int64))
    # a pattern with several components and an isolated vertex
    import scipy.sparse as sp
    A = sp.block_diag([P.laplacian2d_var(10, 1).scipy_full(), sp.identity(1),
                       P.laplacian2d_var(7, 2).scipy_full()]).tocsc()
    low = sp.tril(A).tocsc()
    low.sort_indices()
    perm = P.nd_order(A.shape[0], low.indptr, low.indices)
    assert np.array_equal(np.sort(perm), np.arange(A.shape[0]))
