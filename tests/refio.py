"""Helpers for the reference-analysis harness (oracle/_ref/ref_analyze).

Test infrastructure only: writes the harness input file and parses its record
stream (see oracle/ref_harness.cpp for the layout).
"""
import os
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_analyze")
_KIND = {0: np.int32, 1: np.int64, 2: np.float64}


def have_ref():
    return os.path.exists(REF_BIN)


def write_input(path, n, colptr, rowidx, values, perm):
    with open(path, "wb") as f:
        np.array([n, len(rowidx)], dtype=np.int64).tofile(f)
        np.asarray(colptr, dtype=np.int32).tofile(f)
        np.asarray(rowidx, dtype=np.int32).tofile(f)
        np.asarray(values, dtype=np.float64).tofile(f)
        np.asarray(perm, dtype=np.uint64).tofile(f)


def read_records(path):
    out = {}
    with open(path, "rb") as f:
        while True:
            name = f.read(32)
            if len(name) < 32:
                break
            kind = int(np.frombuffer(f.read(4), dtype=np.int32)[0])
            count = int(np.frombuffer(f.read(8), dtype=np.int64)[0])
            dt = _KIND[kind]
            out[name.rstrip(b"\0").decode()] = np.frombuffer(
                f.read(count * np.dtype(dt).itemsize), dtype=dt).copy()
    return out


def run_ref(prob):
    """Run the reference analysis on a Problem; returns the dumped arrays."""
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        write_input(fin, prob.n, prob.colptr, prob.rowidx, prob.values, prob.perm)
        subprocess.run([REF_BIN, fin, fout], check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        return read_records(fout)
