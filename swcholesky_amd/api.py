"""Host-side interface of the numeric phase, mirroring the reference's entry points.

Reference interface this mirrors (all paths under software/swCholesky/):

* ``analyze_p2`` (cholesky/sw_LSparsity.h:248-702) -> :func:`analyze_p2`, returning
  a :class:`BCSC` with the reference's field names (super, s, i_ptr, p, Perm,
  sParent, col2Sup, nsuper, ssize, xsize) plus prunePtr / pruneSet.
* ``sw_cholesky_left_par_05`` (cholesky/sw_parallel_PB_Cholesky_05.h:52-66) ->
  :func:`sw_cholesky_left_par_05`: same arguments, factors on the GPU, writes
  lValues in place, returns 1 on success and 0 on failure.
* ``blockedLsolve`` (triangularSolve/Triangular_BCSC.h:14-49) and the solve the
  reference lacks -> :class:`Factor` (forward, backward, permutation, refinement).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import SwcError, check, ptr

_SYM = {
    "Perm": (0, "n", np.int32), "ColCount": (1, "n", np.int32), "parent": (2, "n", np.int32),
    "super": (3, "ns1", np.int32), "sParent": (4, "ns", np.int32), "col2Sup": (5, "n", np.int32),
    "pi": (6, "ns1", np.int64), "s": (7, "ssize", np.int32), "i_ptr": (8, "n1", np.int64),
    "p": (9, "n1", np.int64), "prunePtr": (10, "ns1", np.int32),
    "pruneSet": (11, "nprune", np.int32), "a2p": (12, "n1", np.int32),
    "a2i": (13, "nnz", np.int32), "a2src": (14, "nnz", np.int64),
}


class BCSC:
    """Symbolic supernodal factor as analyze_p2 returns it (def.h:123-210 names)."""

    def __init__(self, handle, sizes, flops):
        self._h = handle
        # kept on the instance: a finalizer run while the interpreter shuts down
        # (module globals already cleared) still releases the handle
        self._free = _lib.lib().swc_symbolic_free
        self.n, self.nsuper, self.ssize, self.xsize, self.maxSupWid, self.maxCol, self.nnz, \
            self.nprune = (int(v) for v in sizes)
        self.flops = float(flops)  # sum_j ColCount[j]^2
        dims = {"n": self.n, "n1": self.n + 1, "ns": self.nsuper, "ns1": self.nsuper + 1,
                "ssize": self.ssize, "nprune": self.nprune, "nnz": self.nnz}
        lib = _lib.lib()
        for name, (which, dim, dt) in _SYM.items():
            a = np.zeros(dims[dim], dtype=dt)
            check(lib.swc_symbolic_copy(self._h, which, ptr(a)))
            setattr(self, name, a)

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._free(h)

    def a2_values(self, values):
        """Values of tril(P A P^T) in A2 order from values of A in input order."""
        return np.ascontiguousarray(np.asarray(values, dtype=np.float64)[self.a2src])


def analyze_batch(instances, nrelax=(4, 16, 48), zrelax=(0.8, 0.1, 0.05), threads=0):
    """Symbolic analysis of independent instances [(n, colptr, rowidx, perm), ...]
    on `threads` host threads (0: all cores), concatenated into the analysis of
    their block-diagonal operator (equal to analyze_p2 of that operator)."""
    arrs = [(int(n), np.ascontiguousarray(c, dtype=np.int32), np.ascontiguousarray(r, dtype=np.int32),
             np.ascontiguousarray(p, dtype=np.int64)) for n, c, r, p in instances]
    count = len(arrs)
    ns = np.array([a[0] for a in arrs], dtype=np.int32)
    vp = ctypes.c_void_p * count
    aps = vp(*[a[1].ctypes.data for a in arrs])
    ais = vp(*[a[2].ctypes.data for a in arrs])
    pps = vp(*[a[3].ctypes.data for a in arrs])
    nr = np.asarray(nrelax, dtype=np.int32)
    zr = np.asarray(zrelax, dtype=np.float64)
    h = ctypes.c_void_p()
    lib = _lib.lib()
    check(lib.swc_analyze_batch(count, ptr(ns), ctypes.cast(aps, ctypes.c_void_p),
                                ctypes.cast(ais, ctypes.c_void_p), ctypes.cast(pps, ctypes.c_void_p),
                                ptr(nr), ptr(zr), int(threads), ctypes.byref(h)))
    sizes = np.zeros(8, dtype=np.int64)
    fl = ctypes.c_double()
    check(lib.swc_symbolic_sizes(h, ptr(sizes), ctypes.byref(fl)))
    return BCSC(h, sizes, fl.value)


def analyze_p2(n, colptr, rowidx, perm, nrelax=(4, 16, 48), zrelax=(0.8, 0.1, 0.05)):
    """Symbolic analysis for a GIVEN ordering (perm[new] = old)."""
    colptr = np.ascontiguousarray(colptr, dtype=np.int32)
    rowidx = np.ascontiguousarray(rowidx, dtype=np.int32)
    perm = np.ascontiguousarray(perm, dtype=np.int64)
    nr = np.asarray(nrelax, dtype=np.int32)
    zr = np.asarray(zrelax, dtype=np.float64)
    h = ctypes.c_void_p()
    lib = _lib.lib()
    check(lib.swc_analyze(int(n), ptr(colptr), ptr(rowidx), ptr(perm), ptr(nr), ptr(zr),
                          ctypes.byref(h)))
    sizes = np.zeros(8, dtype=np.int64)
    fl = ctypes.c_double()
    check(lib.swc_symbolic_sizes(h, ptr(sizes), ctypes.byref(fl)))
    return BCSC(h, sizes, fl.value)


def sw_cholesky_left_par_05(n, c, r, values, lC, Ls, Li_ptr, lValues, blockSet, supNo,
                            timing, prunePtr, pruneSet, nLevels=0, levelPtr=None,
                            levelSet=None, nPar=0, parPtr=None, partition=None, chunk=1,
                            threads=1, super_max=0, col_max=0, nodCost=None):
    """The reference numeric entry point on the GPU (host arrays, lValues in place)."""
    arrs = [np.ascontiguousarray(a) for a in (c, r, values, lC, Ls, Li_ptr, blockSet,
                                              prunePtr, pruneSet)]
    c, r, values, lC, Ls, Li_ptr, blockSet, prunePtr, pruneSet = arrs
    assert lValues.dtype == np.float64 and lValues.flags.c_contiguous
    return _lib.lib().sw_cholesky_left_par_05(
        int(n), ptr(c), ptr(r), ptr(values), ptr(lC), ptr(Ls), ptr(Li_ptr), ptr(lValues),
        ptr(blockSet), int(supNo), ptr(timing), ptr(prunePtr), ptr(pruneSet), int(nLevels),
        ptr(levelPtr), ptr(levelSet), int(nPar), ptr(parPtr), ptr(partition), int(chunk),
        int(threads), int(super_max), int(col_max), ptr(nodCost))


def blockedLsolve(n, Lp, Li, Lx, NNZ, Li_ptr, col2sup, sup2col, supNo, x):
    """The reference forward solve L y = x (Triangular_BCSC.h:14-49) on the GPU, in
    place on the host array x; returns 1 on success and 0 on failure."""
    Lp = np.ascontiguousarray(Lp, dtype=np.uint64)
    Li = np.ascontiguousarray(Li, dtype=np.int32)
    Lx = np.ascontiguousarray(Lx, dtype=np.float64)
    Li_ptr = np.ascontiguousarray(Li_ptr, dtype=np.uint64)
    sup2col = np.ascontiguousarray(sup2col, dtype=np.int32)
    col2sup = None if col2sup is None else np.ascontiguousarray(col2sup, dtype=np.int32)
    assert x.dtype == np.float64 and x.flags.c_contiguous and x.size >= n
    return _lib.lib().blockedLsolve(int(n), ptr(Lp), ptr(Li), ptr(Lx), int(NNZ), ptr(Li_ptr),
                                    ptr(col2sup), ptr(sup2col), int(supNo), ptr(x))


def bcsc2csc(bcsc, values):
    """The reference's bcsc2csc (common/Util.h:311-338) on host arrays: returns
    (Cp, Ci, Cx) with 32-bit indices."""
    n = bcsc.n
    nnz = int(sum((bcsc.pi[s + 1] - bcsc.pi[s]) * (bcsc.super[s + 1] - bcsc.super[s])
                  - (bcsc.super[s + 1] - bcsc.super[s]) * (bcsc.super[s + 1] - bcsc.super[s] - 1) // 2
                  for s in range(bcsc.nsuper)))
    Cp = np.zeros(n + 1, dtype=np.int32)
    Ci = np.zeros(max(nnz, 1), dtype=np.int32)
    Cx = np.zeros(max(nnz, 1), dtype=np.float64)
    Ap = np.ascontiguousarray(bcsc.p, dtype=np.uint64)
    AiP = np.ascontiguousarray(bcsc.i_ptr, dtype=np.uint64)
    Ax = np.ascontiguousarray(values, dtype=np.float64)
    ok = _lib.lib().bcsc2csc(n, bcsc.nsuper, ptr(Ap), ptr(bcsc.s), ptr(AiP), ptr(bcsc.super),
                             ptr(Ax), ptr(Cp), ptr(Ci), ptr(Cx))
    if ok != 1:
        raise SwcError(-1, _lib.lib().swc_last_error().decode(errors="replace"))
    return Cp, Ci[:nnz], Cx[:nnz]


def _dptr(x):
    """(pointer, on_device) of a numpy array or a torch tensor."""
    if isinstance(x, np.ndarray):
        return ptr(x), 0
    if hasattr(x, "data_ptr"):
        return ctypes.c_void_p(x.data_ptr()), int(x.is_cuda)
    raise TypeError("expected a numpy array or a torch tensor")


class Factor:
    """GPU numeric factor of one symbolic structure (the handle API).

    ``Factor(bcsc)`` uploads the structure once; :meth:`factor` takes the values of
    A in the analysed input order (host numpy or device torch), :meth:`solve`
    solves A x = b in the original ordering.
    """

    def __init__(self, bcsc: BCSC, params="auto", schedule=None):
        """params: "auto" (the autotuner's table entry for this structure's class,
        swcholesky_amd/tuning.py), None (the library defaults) or a sequence
        {rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow, dag} (swc_set_plan_params;
        dag must be 0: the dependency-driven launch is disabled).
        schedule: None (wavefront levels) or the reference's H-level arrays
        (levelPtr, parPtr, partition) (swc_set_schedule)."""
        self.bcsc = bcsc
        self.n = bcsc.n
        lib = _lib.lib()
        h = ctypes.c_void_p()
        check(lib.swc_create(ctypes.byref(h)))
        self._destroy = lib.swc_destroy  # see BCSC
        self._h = h
        if isinstance(params, str) and params == "auto":
            from .tuning import params_for
            params = params_for(bcsc)
        self.params = None if params is None else tuple(int(v) for v in params)
        if self.params is not None:
            pv = np.asarray(self.params, dtype=np.int32)
            check(lib.swc_set_plan_params(self._h, ptr(pv), len(pv)))
        if schedule is not None:
            lp, pp, part = (np.ascontiguousarray(a, dtype=np.int32) for a in schedule)
            check(lib.swc_set_schedule(self._h, len(lp) - 1, ptr(lp), ptr(pp), ptr(part)))
        check(lib.swc_upload_symbolic(self._h, bcsc._h))
        self.info = None

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._destroy(h)

    def factor(self, values):
        """Numeric factorization; returns LAPACK-style info (0 = success)."""
        if isinstance(values, np.ndarray):
            values = np.ascontiguousarray(values, dtype=np.float64)
        p, dev = _dptr(values)
        info = ctypes.c_int(0)
        rc = _lib.lib().swc_factor(self._h, p, dev, ctypes.byref(info))
        self.info = info.value
        if rc not in (0, -4):
            check(rc)
        return self.info

    def L_values(self):
        out = np.zeros(self.bcsc.xsize, dtype=np.float64)
        check(_lib.lib().swc_get_factor(self._h, ptr(out), 0))
        return out

    def L_csc(self):
        """The factor as a scipy CSC matrix in the layout of the reference's
        bcsc2csc, converted on the GPU (swc_factor_csc)."""
        import scipy.sparse as sp
        lib = _lib.lib()
        nnz = ctypes.c_int64()
        check(lib.swc_factor_nnz(self._h, ctypes.byref(nnz)))
        Cp = np.zeros(self.n + 1, dtype=np.int64)
        Ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        Cx = np.zeros(max(nnz.value, 1), dtype=np.float64)
        check(lib.swc_factor_csc(self._h, ptr(Cp), ptr(Ci), ptr(Cx), 0))
        return sp.csc_matrix((Cx[:nnz.value], Ci[:nnz.value], Cp), shape=(self.n, self.n))

    def solve(self, b):
        """x = A^{-1} b for a vector or an n x nrhs matrix (numpy, returns a copy)."""
        b = np.asarray(b, dtype=np.float64)
        vec = b.ndim == 1
        X = np.asfortranarray(b.reshape(self.n, -1)).copy(order="F")
        check(_lib.lib().swc_solve(self._h, ptr(X), X.shape[1], self.n, 0))
        return X[:, 0].copy() if vec else X

    def solve_device(self, B):
        """In-place solve of a device tensor B (n x nrhs, column-major = B.T contiguous)."""
        p, dev = _dptr(B)
        nrhs = 1 if B.dim() == 1 else B.shape[0]
        check(_lib.lib().swc_solve(self._h, p, nrhs, self.n, dev))
        return B

    def set_precision(self, mode):
        """0: FP64; 1: FP32-MFMA descendant updates; 2: FP32 for every update."""
        check(_lib.lib().swc_set_precision(self._h, int(mode)))

    def refine(self, b, tol=1e-12, maxiter=50, method=-1):
        """x with backward error <= tol by FP64 refinement; returns (x, eta, iters).

        b: numpy vector / n x nrhs matrix, or a device tensor laid out n-major per
        column (torch tensor of shape (nrhs, n) or (n,)), refined in place."""
        lib = _lib.lib()
        if isinstance(b, np.ndarray):
            vec = b.ndim == 1
            B = np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(self.n, -1))
            X = np.zeros_like(B, order="F")
            nrhs = B.shape[1]
            eta = np.zeros(nrhs)
            it = ctypes.c_int(0)
            check(lib.swc_refine(self._h, ptr(B), ptr(X), nrhs, self.n, 0, float(tol),
                                 int(maxiter), int(method), ptr(eta), ctypes.byref(it)))
            return (X[:, 0] if vec else X), eta, it.value
        p, dev = _dptr(b)
        nrhs = 1 if b.dim() == 1 else b.shape[0]
        eta = np.zeros(nrhs)
        it = ctypes.c_int(0)
        check(lib.swc_refine(self._h, p, p, nrhs, self.n, dev, float(tol), int(maxiter),
                             int(method), ptr(eta), ctypes.byref(it)))
        return b, eta, it.value

    def backward_error(self, b, x):
        """eta per column of x for A x = b, computed on the GPU in FP64."""
        B = np.asfortranarray(np.asarray(b, dtype=np.float64).reshape(self.n, -1))
        X = np.asfortranarray(np.asarray(x, dtype=np.float64).reshape(self.n, -1))
        eta = np.zeros(B.shape[1])
        check(_lib.lib().swc_residual(self._h, ptr(B), ptr(X), B.shape[1], self.n, 0, ptr(eta)))
        return eta

    def selected_inverse(self):
        """Z = A^{-1} on the pattern of L (L's BCSC layout, factor ordering)."""
        out = np.zeros(self.bcsc.xsize)
        check(_lib.lib().swc_selected_inverse(self._h, ptr(out), 0))
        return out

    def logdet(self):
        v = ctypes.c_double()
        check(_lib.lib().swc_logdet(self._h, ctypes.byref(v)))
        return v.value

    def gradient(self, x=None, alpha=1.0, beta=0.0, out=None):
        """Gradient of alpha logdet(A) + beta b^T A^{-1} b w.r.t. the stored values
        (x = A^{-1} b).  numpy in -> numpy out; torch cuda tensors in -> `out`."""
        lib = _lib.lib()
        if out is not None and not isinstance(out, np.ndarray):
            xp, xdev = _dptr(x) if x is not None else (None, 1)
            gp, gdev = _dptr(out)
            check(lib.swc_gradient(self._h, xp, xdev, float(alpha), float(beta), gp, gdev))
            return out
        g = np.zeros(self.bcsc.nnz)
        xa = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        check(lib.swc_gradient(self._h, ptr(xa), 0, float(alpha), float(beta), ptr(g), 0))
        return g

    def set_graphs(self, on=True):
        check(_lib.lib().swc_set_graphs(self._h, int(on)))

    def set_profiling(self, on=True):
        check(_lib.lib().swc_set_profiling(self._h, int(on)))

    def timings(self):
        out = np.zeros(8)
        check(_lib.lib().swc_get_timings(self._h, ptr(out)))
        keys = ["factor_ms", "solve_ms", "tile_ms", "tile_launches", "trsm_ms", "trsm_launches",
                "tile_flops", "trsm_flops"]
        return dict(zip(keys, out.tolist()))

    def level_timings(self):
        """(ms, useful flops) of each level's tile-update launch in the last
        profiled factorization, as an nlevel x 2 array."""
        lib = _lib.lib()
        nl = ctypes.c_int(0)
        check(lib.swc_level_timings(self._h, None, 0, ctypes.byref(nl)))
        out = np.zeros((nl.value, 2))
        check(lib.swc_level_timings(self._h, ptr(out), nl.value, ctypes.byref(nl)))
        return out

    def plan_stats(self):
        out = np.zeros(7, dtype=np.int64)
        fl = (ctypes.c_double * 5)()
        check(_lib.lib().swc_plan_stats(self._h, ptr(out), fl))
        return dict(panels=int(out[0]), tiles=int(out[1]), tasks=int(out[2]),
                    levels=int(out[3]), xsize=int(out[4]), dag_items=int(out[5]),
                    dag_level0=int(out[6]), update_flops=fl[0], factor_flops=fl[1],
                    dag_flops=fl[2], dag_sim_ms=fl[3] / 1e3, level_sim_ms=fl[4] / 1e3)


PANEL_DT = np.dtype([("vbase", np.int64), ("voff", np.int64), ("ld", np.int32),
                     ("ncol", np.int32), ("nrow", np.int32), ("c0", np.int32),
                     ("rows_off", np.int32), ("sn", np.int32), ("sc0", np.int32),
                     ("sw", np.int32)])
TILE_DT = np.dtype([("vbase", np.int64), ("panel", np.int32), ("r0", np.int32), ("nr", np.int32),
                    ("task_begin", np.int32), ("task_end", np.int32), ("kind", np.int32),
                    ("slot", np.int32), ("ld", np.int32), ("ncol", np.int32), ("c0", np.int32)])
TASK_DT = np.dtype([("src_base", np.int64), ("ld", np.int32), ("k", np.int32),
                    ("rows_off", np.int32), ("ra", np.int32), ("rb", np.int32),
                    ("ca", np.int32), ("cb", np.int32), ("kind", np.int32),
                    ("rmask", np.uint64), ("cmask", np.uint64)])
SBLK_DT = np.dtype([("panel", np.int32), ("r0", np.int32), ("nr", np.int32),
                    ("writer", np.int32)])
DAG_DT = np.dtype([("tile", np.int32), ("fclass", np.int32), ("prev", np.int32),
                   ("diag", np.int32), ("dep0", np.int32), ("dep1", np.int32)])


def build_plan(bcsc: BCSC, derive_sets=False, params=None, schedule=None):
    """The device work plan for a structure, as numpy record arrays (host only).
    derive_sets: ignore the analysis's descendant sets and derive them from the
    row lists and build the solve part only (what blockedLsolve does); params:
    plan shape parameters as swc_set_plan_params takes them; schedule: the
    H-level arrays (levelPtr, parPtr, partition) as swc_set_schedule takes them."""
    lib = _lib.lib()
    sizes = np.zeros(4, dtype=np.int64)
    check(lib.swc_plan_struct_sizes(ptr(sizes)))
    want = [PANEL_DT.itemsize, TILE_DT.itemsize, TASK_DT.itemsize, SBLK_DT.itemsize]
    if list(sizes) != want:
        raise RuntimeError(f"plan struct layout mismatch: library {list(sizes)}, python {want}")
    h = ctypes.c_void_p()
    super_ = np.ascontiguousarray(bcsc.super)
    p = np.ascontiguousarray(bcsc.p.astype(np.uint64))
    iptr = np.ascontiguousarray(bcsc.i_ptr.astype(np.uint64))
    pp, ps = (None, None) if derive_sets else (bcsc.prunePtr, bcsc.pruneSet)
    pv = None if params is None else np.asarray(params, dtype=np.int32)
    lp = pr = part = None
    if schedule is not None:
        lp, pr, part = (np.ascontiguousarray(a, dtype=np.int32) for a in schedule)
    check(lib.swc_plan_build_params(bcsc.n, ptr(super_), bcsc.nsuper, ptr(p), ptr(bcsc.s),
                                    ptr(iptr), ptr(pp), ptr(ps), ptr(pv),
                                    0 if pv is None else len(pv),
                                    0 if lp is None else len(lp) - 1, ptr(lp), ptr(pr), ptr(part),
                                    ctypes.byref(h)))
    try:
        sz = np.zeros(16, dtype=np.int64)
        fl = (ctypes.c_double * 5)()
        check(lib.swc_plan_sizes(h, ptr(sz), fl))
        (npanel, ntile, ntask, nlevel, n, xsize, ntrsm, nfblk, vsize, nsp, dscr, nslevel, nfused,
         ndag, dag_level0, ndep) = (int(v) for v in sz)
        spec = [(0, PANEL_DT, npanel), (1, TILE_DT, ntile), (2, TASK_DT, ntask),
                (3, np.int32, nlevel + 1), (4, np.int32, nlevel + 1), (5, np.int32, npanel),
                (6, np.int32, nlevel + 1), (7, np.int64, n + 1), (8, np.int64, vsize),
                (9, np.int32, ntrsm), (10, SBLK_DT, nfblk), (11, np.int32, nslevel + 1),
                (12, np.int32, n), (13, np.int32, nslevel + 1), (14, PANEL_DT, nsp),
                (15, np.int32, nsp), (16, np.int32, nslevel + 1),
                (17, np.dtype([("tile0", np.int32), ("ntile", np.int32)]), nfused),
                (18, np.int32, nlevel + 1), (19, np.int32, nlevel), (20, np.int32, nlevel),
                (21, DAG_DT, ndag), (22, np.dtype([("panel", np.int32), ("count", np.int32)]), ndep),
                (23, np.int32, nlevel)]
        names = ["panels", "tiles", "tasks", "lvl_tiles", "lvl_trsm", "panel_order",
                 "lvl_panels", "gptr", "gidx", "trsm_tiles", "fblocks", "lvl_fblk",
                 "gcols", "lvl_gcol", "spanels", "sp_order", "lvl_sp", "fused", "lvl_fused",
                 "lvl_tiles_fused", "lvl_tiles_narrow", "dag_items", "dag_deps",
                 "lvl_tiles_deferred"]
        out = {}
        for (which, dt, cnt), name in zip(spec, names):
            a = np.zeros(cnt, dtype=dt)
            check(lib.swc_plan_export(h, which, ptr(a)))
            out[name] = a
        out.update(nlevel=nlevel, nslevel=nslevel, xsize=xsize, update_flops=fl[0],
                   factor_flops=fl[1], vsize=vsize, diag_scratch=dscr, dag_level0=dag_level0,
                   dag_flops=fl[2], dag_sim_us=fl[3], lvl_sim_us=fl[4])
        return out
    finally:
        lib.swc_plan_free(h)
