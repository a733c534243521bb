"""Batched factorization of many independent SPD instances.

The reference handles one matrix per process (examples/sw_choleskyTest01.cpp).
Here a batch of independent instances is analysed as one block-diagonal
operator: instance k occupies unknowns [off[k], off[k+1]), its own ordering is
kept inside that range, so the elimination forest has one tree per instance.
The instances are analysed independently on host threads and the results
concatenated, which equals the analysis of the block-diagonal operator.
The level schedule then runs level l of every instance in the same launch, and
the whole batch is one captured launch sequence per factorization and per solve.
Each instance's supernodes, descendant sets, tasks and summation order are those
of the instance analysed alone, so its factor is bitwise the same.
"""
from __future__ import annotations

import numpy as np

from . import api


class Batch:
    """A batch of instances analysed and factored together."""

    def __init__(self, problems, nrelax=(4, 16, 48), zrelax=(0.8, 0.1, 0.05), threads=0):
        sizes = np.array([p.n for p in problems], dtype=np.int64)
        self.count = len(problems)
        self.offsets = np.concatenate([[0], np.cumsum(sizes)])
        nnz = np.array([p.nnz for p in problems], dtype=np.int64)
        self.val_offsets = np.concatenate([[0], np.cumsum(nnz)])
        n = int(self.offsets[-1])
        if n > np.iinfo(np.int32).max or self.val_offsets[-1] > np.iinfo(np.int32).max:
            raise ValueError("batch too large for 32-bit indices")
        self.n = n
        self._problems = problems
        self._blockdiag = None
        # the instances are analysed independently on host threads and
        # concatenated (equal to analysing the block-diagonal operator)
        self.bcsc = api.analyze_batch([(p.n, p.colptr, p.rowidx, p.perm) for p in problems],
                                      nrelax, zrelax, threads)
        self._factor = None

    def _assemble(self):
        """(colptr, rowidx, perm) of the block-diagonal operator, built on first use."""
        if self._blockdiag is None:
            n = self.n
            colptr = np.empty(n + 1, dtype=np.int64)
            colptr[0] = 0
            rowidx = np.empty(int(self.val_offsets[-1]), dtype=np.int32)
            perm = np.empty(n, dtype=np.int64)
            for k, p in enumerate(self._problems):
                o, vo = int(self.offsets[k]), int(self.val_offsets[k])
                colptr[o + 1:o + p.n + 1] = p.colptr[1:].astype(np.int64) + vo
                rowidx[vo:vo + p.nnz] = p.rowidx + o
                perm[o:o + p.n] = p.perm + o
            self._blockdiag = (colptr.astype(np.int32), rowidx, perm)
        return self._blockdiag

    @property
    def colptr(self):
        return self._assemble()[0]

    @property
    def rowidx(self):
        return self._assemble()[1]

    @property
    def perm(self):
        return self._assemble()[2]

    @property
    def factor(self):
        """The GPU handle of the whole batch (created on first use)."""
        if self._factor is None:
            self._factor = api.Factor(self.bcsc)
        return self._factor

    def values(self, problems):
        """Concatenated values (input order of every instance)."""
        return np.concatenate([p.values for p in problems])

    def split(self, x):
        """Per-instance views of a concatenated vector (or n x nrhs matrix)."""
        return [x[self.offsets[k]:self.offsets[k + 1]] for k in range(self.count)]

    def instance_factor(self, k, L):
        """L of instance k alone (BCSC arrays re-based to the instance), from the
        batch factor values L (xsize)."""
        b = self.bcsc
        o0, o1 = int(self.offsets[k]), int(self.offsets[k + 1])
        s0 = int(b.col2Sup[o0])
        s1 = int(b.col2Sup[o1 - 1]) + 1
        c0 = int(b.super[s0])
        x0, x1 = int(b.p[c0]), int(b.p[int(b.super[s1])]) if s1 < b.nsuper else int(b.xsize)
        return L[x0:x1], (s0, s1)
