"""Plan shape parameters per structure class, from the autotuner's table.

The table (``tuning.json`` next to this file) is written by ``python -m
swcholesky_amd.tune`` on an MI355X: for each benchmark workload it records the
structure's class (order n, widest supernode) and the parameters
{rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow, dag} that gave the shortest measured
factor + solve time.  A structure takes the entry of the nearest class in
(log2 n, log2 widest supernode); with no table it keeps the library defaults.
"""
from __future__ import annotations

import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "tuning.json")
NAMES = ("rlb", "spw", "narrow_min_tiles", "narrow_max_k", "fuse_narrow", "dag")
DEFAULTS = (8, 128, 1024, 256, 1, 0)

_table = None


def load_table(path=TABLE):
    """The table's entries (a list of dicts), or [] when there is none."""
    global _table
    if path != TABLE:
        with open(path) as fh:
            return json.load(fh)["entries"]
    if _table is None:
        try:
            with open(TABLE) as fh:
                _table = json.load(fh)["entries"]
        except FileNotFoundError:
            _table = []
    return _table


def structure_class(bcsc):
    return {"n": int(bcsc.n), "max_width": int(bcsc.maxSupWid)}


def _distance(a, b):
    return math.hypot(math.log2(max(a["n"], 1)) - math.log2(max(b["n"], 1)),
                      math.log2(max(a["max_width"], 1)) - math.log2(max(b["max_width"], 1)))


def params_for(bcsc, entries=None):
    """(rlb, spw, narrow_min_tiles, narrow_max_k, fuse_narrow, dag) for a symbolic structure."""
    entries = load_table() if entries is None else entries
    if not entries:
        return DEFAULTS
    cls = structure_class(bcsc)
    best = min(entries, key=lambda e: _distance(cls, e["class"]))
    return tuple(int(best["params"].get(k, d)) for k, d in zip(NAMES, DEFAULTS))
