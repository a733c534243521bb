"""Batched instances: the block-diagonal analysis keeps every instance intact."""
import numpy as np
import pytest

import swcholesky_amd as S
from swcholesky_amd import problems as P
from swcholesky_amd.batch import Batch


def _instances(count, seed=3, lo=16, hi=40):
    rng = np.random.default_rng(seed)
    sides = rng.integers(lo, hi + 1, size=count)
    return [P.laplacian2d_var(int(s), seed=None, rng=rng, wlo=1e-1, whi=1e1, shift=1e-4)
            for s in sides]


def test_batch_structure_equals_instances():
    probs = _instances(6)
    bt = Batch(probs)
    B = bt.bcsc
    sn = 0
    for k, p in enumerate(probs):
        b = S.analyze_p2(p.n, p.colptr, p.rowidx, p.perm)
        o = bt.offsets[k]
        assert np.array_equal(B.Perm[o:o + p.n] - o, b.Perm)
        sup = B.super[sn:sn + b.nsuper + 1] - o
        assert np.array_equal(sup, b.super)
        assert np.array_equal(B.s[B.pi[sn]:B.pi[sn + b.nsuper]] - o, b.s)
        pp = B.prunePtr[sn:sn + b.nsuper + 1] - B.prunePtr[sn]
        assert np.array_equal(pp, b.prunePtr)
        ps = B.pruneSet[B.prunePtr[sn]:B.prunePtr[sn + b.nsuper]] - sn
        assert np.array_equal(ps, b.pruneSet)
        sn += b.nsuper
    assert sn == B.nsuper
    assert np.isclose(B.flops, sum(S.analyze_p2(p.n, p.colptr, p.rowidx, p.perm).flops
                                   for p in probs))


@pytest.mark.parametrize("threads", [1, 3, 0])
def test_threaded_batch_analysis_equals_blockdiag_analysis(threads):
    """The instances analysed independently on host threads and concatenated give,
    array for array, the analysis of the block-diagonal operator."""
    probs = _instances(9, seed=5, lo=8, hi=30)
    bt = Batch(probs, threads=threads)
    ref = S.analyze_p2(bt.n, bt.colptr, bt.rowidx, bt.perm)
    got = bt.bcsc
    for f in ("n", "nsuper", "ssize", "xsize", "maxSupWid", "maxCol", "nnz", "nprune", "flops"):
        assert getattr(got, f) == getattr(ref, f), f
    for f in ("Perm", "ColCount", "parent", "super", "sParent", "col2Sup", "pi", "s", "i_ptr",
              "p", "prunePtr", "pruneSet", "a2p", "a2i", "a2src"):
        assert np.array_equal(getattr(got, f), getattr(ref, f)), f
