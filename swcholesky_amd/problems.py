"""Synthetic SPD operators, nested-dissection orderings and the reference's file formats.

The reference reads its matrices from MatrixMarket files (lower triangle, sorted by
column then row, every diagonal present: ``common/Util.h:77-179``) and its
fill-reducing ordering from a binary ``size_t[n]`` file with ``perm[new] = old``
(``examples/sw_choleskyTest01.cpp:70-82``).  The ordering itself is produced
outside the reference (METIS on an x86 machine, ``software/README.md:41-49``),
so this module supplies its own geometric nested dissection for the structured
grids used here and feeds it through the same GIVEN path.

Every generator returns the lower triangle of A as CSC arrays
``(n, colptr:int32[n+1], rowidx:int32[nnz], values:float64[nnz])`` with rows
sorted ascending inside each column and the diagonal first.

Edge-weight conventions (all random streams are ``numpy.random.default_rng(seed)``):
edges are enumerated axis by axis (x-edges, then y, then z), each axis in C order
of the lower endpoint; a weight ``w`` on edge (u,v) adds ``+w`` to A[u,u] and
A[v,v] and ``-w`` to A[u,v].
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

__all__ = [
    "Problem", "laplacian2d_const", "laplacian2d_var", "laplacian3d_var",
    "vector3d_var", "gmrf2d", "dense_spd", "grid_nd_order", "write_mtx", "read_mtx",
    "write_perm", "read_perm", "lower_csc_from_coo", "full_csr_from_lower",
]


class Problem:
    """A symmetric positive definite operator with its grid ordering."""

    def __init__(self, n, colptr, rowidx, values, perm, name, shape=None, dof=1):
        self.n = int(n)
        self.colptr = np.ascontiguousarray(colptr, dtype=np.int32)
        self.rowidx = np.ascontiguousarray(rowidx, dtype=np.int32)
        self.values = np.ascontiguousarray(values, dtype=np.float64)
        self.perm = np.ascontiguousarray(perm, dtype=np.int64)
        self.name = name
        self.shape = shape
        self.dof = dof

    @property
    def nnz(self):
        return int(self.colptr[-1])

    def scipy_lower(self):
        return sp.csc_matrix((self.values, self.rowidx, self.colptr), shape=(self.n, self.n))

    def scipy_full(self):
        low = self.scipy_lower()
        return (low + sp.tril(low, -1).T).tocsr()


def lower_csc_from_coo(n, rows, cols, vals):
    """Sum duplicates, keep the lower triangle, sort rows inside each column."""
    r = np.asarray(rows, dtype=np.int64)
    c = np.asarray(cols, dtype=np.int64)
    v = np.asarray(vals, dtype=np.float64)
    lo = np.maximum(r, c)
    hi = np.minimum(r, c)
    m = sp.coo_matrix((v, (lo, hi)), shape=(n, n)).tocsc()
    m.sum_duplicates()
    m.sort_indices()
    colptr = m.indptr.astype(np.int32)
    rowidx = m.indices.astype(np.int32)
    # every column must hold its diagonal first (readMatrix assumes no empty column)
    first = rowidx[colptr[:-1]]
    if np.any(np.diff(colptr) == 0) or np.any(first != np.arange(n)):
        raise ValueError("operator is missing a diagonal entry")
    return colptr, rowidx, m.data.astype(np.float64)


def full_csr_from_lower(n, colptr, rowidx, values):
    """Both triangles as CSR (the operator the residual is computed with)."""
    low = sp.csc_matrix((values, rowidx, colptr), shape=(n, n))
    full = (low + sp.tril(low, -1).T).tocsr()
    full.sort_indices()
    return (full.indptr.astype(np.int32), full.indices.astype(np.int32),
            full.data.astype(np.float64))


def _grid_edges(shape):
    """Edges of the 2D/3D grid graph, axis by axis, lower endpoint in C order."""
    idx = np.arange(int(np.prod(shape))).reshape(shape)
    us, vs = [], []
    for ax in range(len(shape)):
        sl_lo = [slice(None)] * len(shape)
        sl_hi = [slice(None)] * len(shape)
        sl_lo[ax] = slice(0, shape[ax] - 1)
        sl_hi[ax] = slice(1, shape[ax])
        us.append(idx[tuple(sl_lo)].ravel())
        vs.append(idx[tuple(sl_hi)].ravel())
    return np.concatenate(us), np.concatenate(vs)


def _weighted_laplacian(shape, w, shift):
    n = int(np.prod(shape))
    u, v = _grid_edges(shape)
    deg = np.zeros(n)
    np.add.at(deg, u, w)
    np.add.at(deg, v, w)
    rows = np.concatenate([np.arange(n), v])
    cols = np.concatenate([np.arange(n), u])
    vals = np.concatenate([deg + shift, -w])
    return n, lower_csc_from_coo(n, rows, cols, vals)


def laplacian2d_const(side=100, diag=4.0 + 1e-3, off=-1.0):
    """5-point Laplacian with constant stencil (diag 4+1e-3, off-diagonal -1)."""
    shape = (side, side)
    n = side * side
    u, v = _grid_edges(shape)
    rows = np.concatenate([np.arange(n), v])
    cols = np.concatenate([np.arange(n), u])
    vals = np.concatenate([np.full(n, diag), np.full(u.size, off)])
    cp, ri, va = lower_csc_from_coo(n, rows, cols, vals)
    return Problem(n, cp, ri, va, grid_nd_order(shape), f"lap2d_const_{side}", shape)


def _loguniform(rng, lo, hi, size):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), size))


def laplacian2d_var(side, seed, wlo=1e-1, whi=1e1, shift=1e-4, rng=None):
    """5-point variable-coefficient Laplacian, weights log-uniform in [wlo, whi]."""
    shape = (side, side)
    rng = np.random.default_rng(seed) if rng is None else rng
    u, _ = _grid_edges(shape)
    w = _loguniform(rng, wlo, whi, u.size)
    n, (cp, ri, va) = _weighted_laplacian(shape, w, shift)
    return Problem(n, cp, ri, va, grid_nd_order(shape), f"lap2d_var_{side}", shape)


def laplacian3d_var(side=64, seed=1, wlo=1e-2, whi=1e2, shift=1e-6):
    """7-point variable-coefficient Laplacian on side^3 (SPD, cond ~1e8 at 64^3)."""
    shape = (side, side, side)
    rng = np.random.default_rng(seed)
    u, _ = _grid_edges(shape)
    w = _loguniform(rng, wlo, whi, u.size)
    n, (cp, ri, va) = _weighted_laplacian(shape, w, shift)
    return Problem(n, cp, ri, va, grid_nd_order(shape), f"lap3d_var_{side}", shape)


def vector3d_var(side=48, seed=2, wlo=1e-2, whi=1e2, shift=1e-6):
    """3-dof-per-node operator: each grid edge couples its nodes by w(I + 0.5 u u^T).

    u is a random unit vector per edge; the diagonal 3x3 block of node a is the sum
    of its incident coupling blocks plus shift*I.  Unknown 3a+c is dof c of node a.
    """
    shape = (side, side, side)
    rng = np.random.default_rng(seed)
    a, b = _grid_edges(shape)
    ne = a.size
    w = _loguniform(rng, wlo, whi, ne)
    u = rng.normal(size=(ne, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    blk = w[:, None, None] * (np.eye(3)[None] + 0.5 * u[:, :, None] * u[:, None, :])
    nn = int(np.prod(shape))
    n = 3 * nn
    diag = np.zeros((nn, 3, 3))
    np.add.at(diag, a, blk)
    np.add.at(diag, b, blk)
    diag += shift * np.eye(3)[None]
    ci, cj = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    rows_d = (3 * np.arange(nn)[:, None, None] + ci[None]).ravel()
    cols_d = (3 * np.arange(nn)[:, None, None] + cj[None]).ravel()
    rows_o = (3 * b[:, None, None] + ci[None]).ravel()
    cols_o = (3 * a[:, None, None] + cj[None]).ravel()
    rows = np.concatenate([rows_d, rows_o])
    cols = np.concatenate([cols_d, cols_o])
    vals = np.concatenate([diag.ravel(), -blk.ravel()])
    keep = rows >= cols
    cp, ri, va = lower_csc_from_coo(n, rows[keep], cols[keep], vals[keep])
    return Problem(n, cp, ri, va, grid_nd_order(shape, dof=3), f"vec3d_var_{side}",
                   shape, dof=3)


def dense_spd(n=1100, seed=3):
    """Dense SPD matrix M M^T / n + I (M uniform in [-1, 1)) in the identity
    ordering: a single supernode of n columns, factored entirely by the
    right-looking steps of a wide supernode."""
    rng = np.random.default_rng(seed)
    m = rng.uniform(-1.0, 1.0, (n, n))
    a = m @ m.T / n + np.eye(n)
    r, c = np.tril_indices(n)
    cp, ri, va = lower_csc_from_coo(n, r, c, a[r, c])
    return Problem(n, cp, ri, va, np.arange(n), f"dense_{n}")


def gmrf2d(side=256, seed=4, tau=1.0, kappa2=0.05, wlo=0.5, whi=2.0):
    """GMRF precision Q = tau (kappa^2 I + L_grid), L_grid weights log-uniform."""
    shape = (side, side)
    rng = np.random.default_rng(seed)
    u, _ = _grid_edges(shape)
    w = _loguniform(rng, wlo, whi, u.size)
    n, (cp, ri, va) = _weighted_laplacian(shape, w, 0.0)
    va = va.copy()
    diag_pos = cp[:-1]
    va[diag_pos] += kappa2
    va *= tau
    return Problem(n, cp, ri, va, grid_nd_order(shape), f"gmrf2d_{side}", shape)


# ---------------------------------------------------------------------------
# geometric nested dissection
# ---------------------------------------------------------------------------

def grid_nd_order(shape, dof=1, leaf=8):
    """Nested-dissection ordering of a structured grid, ``perm[new] = old``.

    The longest axis of each box is cut in the middle by a one-node-thick plane;
    the two halves are ordered first (recursively), the separator last.  Boxes
    with at most ``leaf`` nodes keep their natural (C) order.  With ``dof`` > 1
    the unknowns of one node stay contiguous.
    """
    shape = tuple(int(s) for s in shape)
    idx = np.arange(int(np.prod(shape))).reshape(shape)
    out = []

    def rec(box):
        sub = idx[tuple(slice(lo, hi) for lo, hi in box)]
        if sub.size <= leaf:
            out.append(sub.ravel())
            return
        lens = [hi - lo for lo, hi in box]
        ax = int(np.argmax(lens))
        lo, hi = box[ax]
        mid = lo + (hi - lo) // 2
        left = list(box); left[ax] = (lo, mid)
        right = list(box); right[ax] = (mid + 1, hi)
        sep = list(box); sep[ax] = (mid, mid + 1)
        if mid > lo:
            rec(left)
        if hi > mid + 1:
            rec(right)
        out.append(idx[tuple(slice(a, b) for a, b in sep)].ravel())

    rec([(0, s) for s in shape])
    nodes = np.concatenate(out).astype(np.int64)
    assert nodes.size == idx.size
    if dof == 1:
        return nodes
    return (dof * nodes[:, None] + np.arange(dof)[None, :]).ravel()


# ---------------------------------------------------------------------------
# reference file formats
# ---------------------------------------------------------------------------

def nd_order(n, colptr, rowidx):
    """Nested-dissection ordering of a general symmetric pattern (lower CSC),
    ``perm[new] = old``: the library's own graph heuristic (swc_nd_order), for
    operators without grid coordinates."""
    import ctypes
    from . import _lib
    colptr = np.ascontiguousarray(colptr, dtype=np.int32)
    rowidx = np.ascontiguousarray(rowidx, dtype=np.int32)
    perm = np.zeros(int(n), dtype=np.int64)
    _lib.check(_lib.lib().swc_nd_order(int(n), ctypes.c_void_p(colptr.ctypes.data),
                                       ctypes.c_void_p(rowidx.ctypes.data),
                                       ctypes.c_void_p(perm.ctypes.data)))
    return perm


def mesh2d_random(npts=20000, seed=5, wlo=1e-1, whi=1e1, shift=1e-4):
    """Variable-coefficient Laplacian on an unstructured 2D mesh: the Delaunay
    triangulation of `npts` uniform random points in the unit square, edge
    weights log-uniform in [wlo, whi], diagonal = sum of incident weights +
    shift; ordered by nd_order (no grid to cut along)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = rng.uniform(0, 1, (npts, 2))
    tri = Delaunay(pts).simplices
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [0, 2]]])
    e = np.unique(np.sort(e, axis=1), axis=0)
    w = _loguniform(rng, wlo, whi, len(e))
    diag = np.full(npts, shift)
    np.add.at(diag, e[:, 0], w)
    np.add.at(diag, e[:, 1], w)
    rows = np.concatenate([e[:, 1], np.arange(npts)])
    cols = np.concatenate([e[:, 0], np.arange(npts)])
    vals = np.concatenate([-w, diag])
    colptr, rowidx, values = lower_csc_from_coo(npts, rows, cols, vals)
    perm = nd_order(npts, colptr, rowidx)
    return Problem(npts, colptr, rowidx, values, perm, f"mesh2d_{npts}")


def write_mtx(path, n, colptr, rowidx, values):
    """MatrixMarket lower triangle sorted by (col,row), as ``readMatrix`` expects."""
    cols = np.repeat(np.arange(n), np.diff(colptr))
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{n} {n} {len(rowidx)}\n")
        data = np.column_stack([rowidx + 1, cols + 1, values])
        np.savetxt(f, data, fmt=["%d", "%d", "%.17g"])


def read_mtx(path):
    """Read a lower-triangle MatrixMarket file into sorted lower CSC arrays."""
    with open(path) as f:
        header = f.readline().lower().split()
        if len(header) < 5 or header[0] != "%%matrixmarket" or header[2] != "coordinate":
            raise ValueError("not a coordinate MatrixMarket file")
        if header[3] != "real":
            raise ValueError("only real matrices are supported")
        line = f.readline()
        while line.startswith("%"):
            line = f.readline()
        n, _, nnz = (int(t) for t in line.split())
        data = np.loadtxt(f, ndmin=2)
    if data.shape[0] != nnz:
        raise ValueError("entry count does not match the header")
    r = data[:, 0].astype(np.int64) - 1
    c = data[:, 1].astype(np.int64) - 1
    return (n,) + lower_csc_from_coo(n, r, c, data[:, 2])


def write_perm(path, perm):
    np.asarray(perm, dtype=np.uint64).tofile(path)


def read_perm(path, n):
    p = np.fromfile(path, dtype=np.uint64, count=n).astype(np.int64)
    if p.size != n:
        raise ValueError("short permutation file")
    return p
