"""The host analysis against the reference's own analyze_p2 (golden + live)."""
import os

import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
import refio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CASES = {
    "c1_lap2d_const_100": lambda: P.laplacian2d_const(100),
    "lap2d_var_37": lambda: P.laplacian2d_var(37, 3),
    "lap3d_var_12": lambda: P.laplacian3d_var(12),
    "vec3d_var_6": lambda: P.vector3d_var(6),
    "gmrf2d_32": lambda: P.gmrf2d(32),
}
# reference record name -> BCSC attribute
MAP = {"perm": "Perm", "colcount": "ColCount", "super": "super", "sparent": "sParent",
       "col2sup": "col2Sup", "pi": "pi", "s": "s", "i_ptr": "i_ptr", "p": "p",
       "pruneptr": "prunePtr", "pruneset": "pruneSet", "a2_p": "a2p", "a2_i": "a2i"}


def _compare(ref, prob):
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    hdr = np.array([b.n, b.nsuper, b.ssize, b.xsize, b.maxSupWid, b.maxCol])
    assert np.array_equal(hdr, ref["header"])
    for k, attr in MAP.items():
        assert np.array_equal(getattr(b, attr), ref[k]), k
    # A2 values: same entries as the reference's ptranspose pair
    assert np.array_equal(b.a2_values(prob.values), ref["a2_x"])
    return b


@pytest.mark.parametrize("name", list(CASES))
def test_analysis_matches_golden(name):
    ref = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    _compare(ref, CASES[name]())


@pytest.mark.skipif(not refio.have_ref(), reason="reference harness not built")
@pytest.mark.parametrize("make", [
    lambda: P.laplacian2d_const(10), lambda: P.laplacian2d_var(64, 7),
    lambda: P.laplacian3d_var(24), lambda: P.vector3d_var(9), lambda: P.gmrf2d(100),
    lambda: P.laplacian2d_var(128, 3), lambda: P.laplacian2d_var(16, 3)])
def test_analysis_matches_live_reference(make):
    prob = make()
    _compare(refio.run_ref(prob), prob)


def test_invariants_and_closed_forms():
    """What the mathematics fixes, independently of the reference build."""
    prob = P.laplacian3d_var(8)
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    n = b.n
    assert sorted(b.Perm) == list(range(n))
    # brute-force symbolic Cholesky of P A P^T
    A = prob.scipy_full().toarray()[np.ix_(b.Perm, b.Perm)] != 0
    Lpat = np.tril(A).copy()
    for j in range(n):
        rows = np.nonzero(Lpat[j + 1:, j])[0] + j + 1
        Lpat[np.ix_(rows, rows)] |= True
    Lpat = np.tril(Lpat)
    colcount = Lpat.sum(axis=0)
    assert np.array_equal(colcount, b.ColCount)
    assert np.isclose(b.flops, float((colcount.astype(np.float64) ** 2).sum()))
    parent = np.array([(np.nonzero(Lpat[j + 1:, j])[0][0] + j + 1) if Lpat[j + 1:, j].any()
                       else -1 for j in range(n)])
    assert np.array_equal(parent, b.parent)
    # supernodes partition the columns; each row list covers its columns' patterns
    assert b.super[0] == 0 and b.super[-1] == n and np.all(np.diff(b.super) > 0)
    for s in range(b.nsuper):
        c0, c1 = b.super[s], b.super[s + 1]
        rows = b.s[b.pi[s]:b.pi[s + 1]]
        assert np.all(np.diff(rows) > 0) and np.array_equal(rows[:c1 - c0], np.arange(c0, c1))
        for c in range(c0, c1):
            assert set(np.nonzero(Lpat[:, c])[0]) <= set(rows)
    # pruned descendants of s are exactly the supernodes with a row inside s
    for s in range(b.nsuper):
        c0, c1 = b.super[s], b.super[s + 1]
        want = {d for d in range(s) if np.any((b.s[b.pi[d]:b.pi[d + 1]] >= c0) &
                                              (b.s[b.pi[d]:b.pi[d + 1]] < c1))}
        got = b.pruneSet[b.prunePtr[s]:b.prunePtr[s + 1]]
        assert len(got) == len(set(got)) and set(got) == want


def test_bad_perm_rejected():
    prob = P.laplacian2d_const(5)
    perm = prob.perm.copy()
    perm[0] = perm[1]
    with pytest.raises(S.SwcError):
        S.analyze_p2(prob.n, prob.colptr, prob.rowidx, perm)


@pytest.mark.parametrize("name", ["lap3d_var_64", "vec3d_var_48"])
def test_full_size_analysis_matches_reference_digests(name):
    """At full size (n = 262,144 and 331,776) the restated analysis equals the
    reference's analyze_p2 array for array: SHA-256 digests of the reference's
    outputs, written by tests/golden/make_golden.py from oracle/_ref."""
    import hashlib
    import os
    import swcholesky_amd as S
    from swcholesky_amd import problems as P
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"full_{name}.npz"))
    prob = {"lap3d_var_64": lambda: P.laplacian3d_var(64), "vec3d_var_48": lambda: P.vector3d_var(48)}[name]()
    b = S.analyze_p2(prob.n, prob.colptr, prob.rowidx, prob.perm)
    ours = {"perm": b.Perm, "colcount": b.ColCount, "super": b.super, "s": b.s,
            "pruneptr": b.prunePtr, "pruneset": b.pruneSet, "p": b.p, "i_ptr": b.i_ptr}
    for key, dig, dt in zip(g["digest_keys"], g["digests"], g["dtypes"]):
        a = np.ascontiguousarray(ours[str(key)].astype(str(dt)))
        assert hashlib.sha256(a.tobytes()).hexdigest() == str(dig), key
    assert int(g["header"][1]) == b.nsuper
