"""GPU parity: the HIP factorization and solves against the CPU oracle."""
import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
import oracle_py as O
from helpers import SMALL, full_csr, relF, setup

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,make", SMALL, ids=[s[0] for s in SMALL])
def test_factor_matches_oracle(name, make):
    prob = make()
    b, a2x, Lo, info = setup(prob)
    assert info == 0
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    Lg = f.L_values()
    # FP64 L within 1e-12 relative Frobenius of the oracle factor
    assert relF(Lg, Lo) < 1e-12, relF(Lg, Lo)


@pytest.mark.parametrize("name,make", SMALL, ids=[s[0] for s in SMALL])
def test_solve_backward_error(name, make):
    prob = make()
    b, a2x, Lo, info = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    rng = np.random.default_rng(0)
    xs = rng.uniform(-1, 1, prob.n)
    rhs = A @ xs
    x = f.solve(rhs)
    eta = O.backward_error(rp, ci, v, x, rhs)
    assert eta < 1e-14, eta
    xo = O.solve(b, Lo, rhs)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-6


def test_multi_rhs():
    prob = SMALL[2][1]()
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    B = np.random.default_rng(1).uniform(-1, 1, (prob.n, 8))
    X = f.solve(B)
    for j in range(8):
        assert O.backward_error(rp, ci, v, X[:, j], B[:, j]) < 1e-14


def test_legacy_entry_point_matches_oracle():
    prob = SMALL[0][1]()
    b, a2x, Lo, _ = setup(prob)
    lv = np.zeros(b.xsize)
    timing = np.zeros(8)
    ok = S.sw_cholesky_left_par_05(b.n, b.a2p, b.a2i, a2x, b.p.astype(np.uint64), b.s,
                                   b.i_ptr.astype(np.uint64), lv, b.super, b.nsuper, timing,
                                   b.prunePtr, b.pruneSet)
    assert ok == 1
    assert relF(lv, Lo) < 1e-12
    # strict upper part of every diagonal block stays zero
    for s in range(b.nsuper):
        c0, c1 = b.super[s], b.super[s + 1]
        R = b.pi[s + 1] - b.pi[s]
        blk = lv[b.p[c0]:b.p[c0] + R * (c1 - c0)].reshape(c1 - c0, R).T
        assert np.all(np.triu(blk[:c1 - c0, :], 1) == 0)


def test_not_positive_definite_reports_column():
    from swcholesky_amd import problems as P
    prob = P.laplacian2d_var(20, 5, shift=-2.0)  # strongly indefinite
    b, a2x, Lo, info_o = setup(prob)
    assert info_o > 0
    f = S.Factor(b)
    info = f.factor(prob.values)
    assert info == info_o
    with pytest.raises(S.SwcError):
        f.solve(np.ones(prob.n))


def test_refactor_same_structure_new_values():
    prob = SMALL[3][1]()
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    v2 = prob.values * 2.0
    assert f.factor(v2) == 0
    Lo2, _ = O.factor(b, b.a2_values(v2))
    assert relF(f.L_values(), Lo2) < 1e-12


def test_graph_replays_match_eager_launches():
    """Captured factor/solve sequences replayed several times, with eager work
    and synchronizations in between, give bitwise the eager results."""
    from swcholesky_amd import problems as P
    prob = P.laplacian2d_const(100)
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b)
    f.set_graphs(False)
    assert f.factor(prob.values) == 0
    L_eager = f.L_values()
    rhs = np.random.default_rng(5).uniform(-1, 1, (prob.n, 3))
    x_eager = f.solve(rhs)
    f.set_graphs(True)
    for rep in range(4):
        vals = prob.values * (1.0 + 0.5 * rep)
        assert f.factor(vals) == 0
        f.set_graphs(False)  # eager work between replays
        Lo_rep, _ = O.factor(b, b.a2_values(vals))
        f.set_graphs(True)
        assert relF(f.L_values(), Lo_rep) < 1e-12
    assert f.factor(prob.values) == 0
    assert np.array_equal(f.L_values(), L_eager)
    for rep in range(3):
        assert np.array_equal(f.solve(rhs), x_eager)
        _ = f.solve(rhs[:, :1])  # a different shape in between
    for mode in (1, 0):  # precision switches reuse their own captured sequence
        f.set_precision(mode)
        assert f.factor(prob.values) == 0
    assert np.array_equal(f.L_values(), L_eager)


def test_reupload_larger_structure_into_one_handle():
    """A handle that solved a small structure is reused for a larger one: the
    solve buffers sized by the structure must follow it."""
    small, large = SMALL[0][1](), SMALL[3][1]()
    bs = S.analyze_p2(small.n, small.colptr, small.rowidx, small.perm)
    bl = S.analyze_p2(large.n, large.colptr, large.rowidx, large.perm)
    f = S.Factor(bs)
    assert f.factor(small.values) == 0
    f.solve(np.ones(small.n))
    S._lib.check(S._lib.lib().swc_upload_symbolic(f._h, bl._h))
    f.bcsc, f.n = bl, bl.n
    assert f.factor(large.values) == 0
    A, rp, ci, v = full_csr(large)
    rhs = A @ np.random.default_rng(3).uniform(-1, 1, large.n)
    x = f.solve(rhs)
    assert O.backward_error(rp, ci, v, x, rhs) < 1e-14


@pytest.mark.parametrize("n", [17, 64, 65, 129, 192, 193, 300])
def test_single_supernode_widths(n):
    """A dense SPD matrix is one supernode of n columns: its panels and solve panels
    take every width class (one 64-column chunk, two chunks with a one-column second
    chunk, a trailing narrow panel, ...).  Factor and two right-hand sides against the
    oracle."""
    prob = P.dense_spd(n, seed=n)
    b, a2x, Lo, info = setup(prob)
    assert info == 0 and b.nsuper == 1
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    assert relF(f.L_values(), Lo) < 1e-12
    A, rp, ci, v = full_csr(prob)
    X0 = np.random.default_rng(n).uniform(-1, 1, (n, 2))
    B = A @ X0
    X = f.solve(B)
    for k in range(2):
        assert O.backward_error(rp, ci, v, X[:, k], B[:, k]) < 1e-14
    assert np.linalg.norm(X - X0) / np.linalg.norm(X0) < 1e-10


@pytest.mark.parametrize("name,make", [SMALL[3], SMALL[4], SMALL[6]],
                         ids=[SMALL[3][0], SMALL[4][0], SMALL[6][0]])
@pytest.mark.parametrize("nrhs", [3, 4, 5, 9, 17])
def test_rhs_groups_match_single_rhs(name, make, nrhs):
    """Several right-hand sides are solved in groups (one read of L per group); a
    last group may be partly empty.  Each column of a multi-RHS solve equals the
    single-RHS solve of that column, and reaches the oracle's backward error."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    A, rp, ci, v = full_csr(prob)
    B = np.random.default_rng(nrhs).uniform(-1, 1, (prob.n, nrhs))
    X = f.solve(B)
    for j in range(nrhs):
        xj = f.solve(B[:, j])
        assert np.allclose(X[:, j], xj, rtol=1e-13, atol=1e-13 * np.abs(xj).max())
        assert O.backward_error(rp, ci, v, X[:, j], B[:, j]) < 1e-14


def test_refactor_after_multi_rhs_solve_uses_new_norm():
    """A solve with 8 right-hand sides grows the solve workspace after the factor
    graph was captured; the replayed refactorization must still report ||A||_inf of
    the new values: backward errors of a perturbed solution agree with the oracle's
    (which sees only the new values) at rtol 1e-10."""
    prob = SMALL[3][1]()
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    B = np.random.default_rng(4).uniform(-1, 1, (prob.n, 8))
    f.solve(B)
    v2 = prob.values * 2.0
    assert f.factor(v2) == 0
    Lo2, _ = O.factor(b, b.a2_values(v2))
    assert relF(f.L_values(), Lo2) < 1e-12
    prob2 = P.Problem(prob.n, prob.colptr, prob.rowidx, v2, prob.perm, "x2")
    A, rp, ci, v = full_csr(prob2)
    x, eta, iters = f.refine(B[:, 0], tol=1e-12)
    assert eta[0] <= 1e-12 and O.backward_error(rp, ci, v, x, B[:, 0]) <= 1e-12
    xp = x + 1e-6 * np.random.default_rng(5).standard_normal(prob.n)
    assert np.isclose(f.backward_error(B[:, 0], xp)[0], O.backward_error(rp, ci, v, xp, B[:, 0]),
                      rtol=1e-10)


def test_mixed_after_reupload_of_larger_structure():
    """mixed -> FP64 -> upload of a larger structure -> mixed: the FP32 copy of L
    follows the structure, and the mixed factor is the oracle's FP32-update factor
    up to FP32 rounding."""
    small, large = SMALL[0][1](), SMALL[3][1]()
    bs = S.analyze_p2(small.n, small.colptr, small.rowidx, small.perm)
    bl = S.analyze_p2(large.n, large.colptr, large.rowidx, large.perm)
    assert bl.xsize > bs.xsize
    f = S.Factor(bs)
    f.set_precision(1)
    assert f.factor(small.values) == 0
    f.set_precision(0)
    assert f.factor(small.values) == 0
    S._lib.check(S._lib.lib().swc_upload_symbolic(f._h, bl._h))
    f.bcsc, f.n = bl, bl.n
    f.set_precision(1)
    assert f.factor(large.values) == 0
    a2x = bl.a2_values(large.values)
    L64, _ = O.factor(bl, a2x)
    L32, _ = O.factor(bl, a2x, fp32_updates=True)
    Lm = f.L_values()
    assert relF(Lm, L32) < 10 * max(relF(L32, L64), 1e-7)
    A, rp, ci, v = full_csr(large)
    rhs = A @ np.random.default_rng(6).uniform(-1, 1, large.n)
    x, eta, _ = f.refine(rhs, tol=1e-12)
    assert eta[0] <= 1e-12 and O.backward_error(rp, ci, v, x, rhs) <= 1e-12


@pytest.mark.parametrize("name,make", [SMALL[0], SMALL[3], SMALL[4], SMALL[6]],
                         ids=[SMALL[0][0], SMALL[3][0], SMALL[4][0], SMALL[6][0]])
def test_blockedLsolve_drop_in(name, make):
    """The reference's forward-solve symbol, called with the arguments
    triangularTest02.cpp:195 passes, against the oracle's forward solve, and the
    closed form b = L 1 => y = 1 of rhsInitBlocked (Util.h:277-306).  A second call
    on the same structure reuses the kept schedule.  The closed form's forward
    error grows with the conditioning of L (the oracle itself is 6e-12 from 1 on
    lap3d_var_20), so it is held to the oracle's own error, and the residual
    ||L y - b|| / (||L|| ||y||) to 1e-14 (the oracle reaches 1.3e-15)."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    for rep in range(2):
        x = np.random.default_rng(rep).uniform(-1, 1, b.n)
        yo = O.lsolve(b, Lo, x.copy())[:, 0]
        y = x.copy()
        ok = S.blockedLsolve(b.n, b.p, b.s, Lo, b.nnz, b.i_ptr, b.col2Sup, b.super, b.nsuper, y)
        assert ok == 1
        assert np.linalg.norm(y - yo) / np.linalg.norm(yo) <= 1e-14
    Ld = O.to_sparse_L(b, Lo)
    rhs = Ld @ np.ones(b.n)
    err_o = np.max(np.abs(O.lsolve(b, Lo, rhs.copy())[:, 0] - 1.0))
    y = rhs.copy()
    assert S.blockedLsolve(b.n, b.p, b.s, Lo, b.nnz, b.i_ptr, None, b.super, b.nsuper, y) == 1
    assert np.max(np.abs(y - 1.0)) <= 4 * err_o + 1e-14
    res = np.max(np.abs(Ld @ y - rhs)) / (abs(Ld).sum(axis=1).max() * np.max(np.abs(y)))
    assert res <= 1e-14


@pytest.mark.parametrize("name", ["c1_lap2d_const_100", "lap3d_var_12", "vec3d_var_6"])
def test_hlevel_schedule_matches_oracle(name):
    """Factor levels from the reference's own H-level schedule (golden fixture
    of its analyze_p2): the same factor as the oracle, and as the wavefront
    schedule's within rounding."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    prob = {"c1_lap2d_const_100": lambda: P.laplacian2d_const(100),
            "lap3d_var_12": lambda: P.laplacian3d_var(12),
            "vec3d_var_6": lambda: P.vector3d_var(6)}[name]()
    b, a2x, Lo, _ = setup(prob)
    f = S.Factor(b, schedule=(g["levelptr"], g["parptr"], g["partition"]))
    assert f.factor(prob.values) == 0
    assert relF(f.L_values(), Lo) <= 1e-12
    A, rp, ci, v = full_csr(prob)
    rhs = A @ np.random.default_rng(8).uniform(-1, 1, prob.n)
    x = f.solve(rhs)
    assert O.backward_error(rp, ci, v, x, rhs) <= 1e-14


def test_factor_from_mtx_and_perm_files_exported_as_csc(tmp_path):
    """The reference's file path end to end: the operator written as
    MatrixMarket lower triangle and a binary size_t perm file, read back,
    analysed and factored on the GPU, the factor exported to CSC on the GPU
    (swc_factor_csc) and compared with the host bcsc2csc of the oracle's factor:
    same pattern, values within 1e-12."""
    from swcholesky_amd import api
    prob = P.laplacian3d_var(20)
    P.write_mtx(str(tmp_path / "a.mtx"), prob.n, prob.colptr, prob.rowidx, prob.values)
    P.write_perm(str(tmp_path / "a.perm"), prob.perm)
    n, cp, ri, v = P.read_mtx(str(tmp_path / "a.mtx"))
    perm = P.read_perm(str(tmp_path / "a.perm"), n)
    b = S.analyze_p2(n, cp, ri, perm)
    f = S.Factor(b)
    assert f.factor(v) == 0
    Lc = f.L_csc()
    Lo, info = O.factor(b, b.a2_values(v))
    assert info == 0
    Cp, Ci, Cx = api.bcsc2csc(b, Lo)
    assert np.array_equal(Lc.indptr, Cp) and np.array_equal(Lc.indices, Ci)
    assert np.linalg.norm(Lc.data - Cx) <= 1e-12 * np.linalg.norm(Cx)


def test_unstructured_mesh_with_own_nd_order():
    """An operator without grid coordinates (Delaunay mesh of 20,000 random
    points) ordered by the library's nested dissection: GPU factor within 1e-12
    of the oracle's, solve to backward error 1e-14."""
    prob = P.mesh2d_random(20000)
    b, a2x, Lo, info = setup(prob)
    assert info == 0
    f = S.Factor(b)
    assert f.factor(prob.values) == 0
    assert relF(f.L_values(), Lo) <= 1e-12
    A, rp, ci, v = full_csr(prob)
    rhs = A @ np.random.default_rng(9).uniform(-1, 1, prob.n)
    x = f.solve(rhs)
    assert O.backward_error(rp, ci, v, x, rhs) <= 1e-14


@pytest.mark.parametrize("name,make", [SMALL[0], SMALL[3], SMALL[5]],
                         ids=[SMALL[0][0], SMALL[3][0], SMALL[5][0]])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_narrow_panels_match_two_kernel_path(name, make, mode):
    """Narrow panels updated and factored by the persistent worklist kernel (one
    wave per panel) give bitwise the factor of the update + factor launches they
    replace, and the oracle's within 1e-12 (FP64); forced on small operators with
    a one-tile threshold."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    out = {}
    for fuse in (1, 0):
        f = S.Factor(b, params=(8, 128, 1, 256, fuse))
        f.set_precision(mode)
        assert f.factor(prob.values) == 0
        out[fuse] = f.L_values()
    assert np.array_equal(out[0], out[1])
    if mode == 0:
        assert relF(out[1], Lo) <= 1e-12


DAG_CASES = [SMALL[2], SMALL[3], SMALL[4], SMALL[6], ("lap3d_var_32", lambda: P.laplacian3d_var(32))]

# The dependency-driven launch is switched off (swc_set_plan_params rejects dag != 0):
# one run of the test below gave a non-positive pivot (info 32549 on lap3d_var_32,
# mode 1, order 1) where the per-level launches give none, and no cause was found in
# the launch's waits and fences.
DAG_OFF = "dependency-driven launch disabled: an intermittent wrong factor, cause not found"


@pytest.mark.skip(reason=DAG_OFF)
@pytest.mark.parametrize("name,make", DAG_CASES, ids=[c[0] for c in DAG_CASES])
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_dependency_driven_launch_matches_level_launches(name, make, mode):
    """The levels after the narrow-panel ones as one dependency-driven launch (in
    list-schedule order, in level order, and with every level in it) give bitwise
    the factor of the per-level launches, on the first (captured) and later
    (replayed) factorizations, and the oracle's within 1e-12 in FP64."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    ref = S.Factor(b, params=(8, 128, 1024, 256, 1, 0))
    ref.set_precision(mode)
    assert ref.factor(prob.values) == 0
    Lref = ref.L_values()
    if mode == 0:
        assert relF(Lref, Lo) <= 1e-12
    for params in ((8, 128, 1024, 256, 1, 1), (8, 128, 1024, 256, 1, 2),
                   (8, 128, 1 << 30, 256, 0, 1)):
        print(f"[dag] {name} mode {mode} params {params}", flush=True)
        f = S.Factor(b, params=params)
        f.set_precision(mode)
        for rep in range(3):
            assert f.factor(prob.values) == 0
            assert np.array_equal(f.L_values(), Lref), (params, rep)


@pytest.mark.skip(reason=DAG_OFF)
def test_dependency_driven_launch_reports_non_positive_pivot():
    prob = P.laplacian2d_var(20, 5, shift=-2.0)  # strongly indefinite
    b, a2x, Lo, info_o = setup(prob)
    assert info_o > 0
    for params in ((8, 128, 1 << 30, 256, 0, 1), (8, 128, 1 << 30, 256, 0, 0)):
        f = S.Factor(b, params=params)
        assert f.factor(prob.values) == info_o


def test_dependency_driven_launch_is_rejected():
    """swc_set_plan_params refuses dag = 1 or 2 with an error naming the launch."""
    prob = P.laplacian3d_var(8)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    for dag in (1, 2):
        with pytest.raises(Exception, match="dependency-driven"):
            S.Factor(b, params=(8, 128, 1024, 256, 1, dag))


@pytest.mark.parametrize("name,make", [("dense_2000", lambda: P.dense_spd(2000))],
                         ids=["dense_2000"])
def test_deferred_updates_beside_panel_factor_match_separate_launches(name, make):
    """Levels that run their other panels' updates in the launch of their panel
    factorization give bitwise the factor of separate update and factor launches
    (what the per-launch timing mode runs)."""
    prob = make()
    b, a2x, Lo, _ = setup(prob)
    assert S.api.build_plan(b)["lvl_tiles_deferred"].sum() > 0
    out = []
    for profile in (False, True):
        f = S.Factor(b, params=None)
        f.set_profiling(profile)
        assert f.factor(prob.values) == 0
        out.append(f.L_values())
    assert np.array_equal(out[0], out[1])
    assert relF(out[0], Lo) <= 1e-12
