"""Builds libswcholesky_amd.so in-tree with hipcc for gfx950.

    python -m swcholesky_amd.build

Host C++ (analysis, C ABI glue) and HIP device code go into one shared library
next to this file; Python loads it with ctypes (swcholesky_amd/_lib.py).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libswcholesky_amd.so")

HOST_SOURCES = ["analysis.cpp", "capi_analysis.cpp", "capi_common.cpp"]
HIP_SOURCES = ["numeric.hip"]
HOST_SOURCES += ["plan.cpp", "ordering.cpp"]

# -packed-fp32-ops: keep hipcc from forming packed FP32 instructions (v_pk_*_f32),
# some forms of which have given wrong results on this hardware.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    objs = []
    odir = os.path.join(HERE, "_obj")
    os.makedirs(odir, exist_ok=True)
    for src in HOST_SOURCES + HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(odir, src + ".o")
        if src.endswith(".hip"):
            cmd = ["hipcc", *HIP_FLAGS, "-c", path, "-o", obj]
        else:
            cmd = ["hipcc", "-O3", "-std=c++17", "-fPIC", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
