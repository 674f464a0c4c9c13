"""Reads a core's quadtree through the ABI (tree_info / tree_order) into the
trace format of the reference's scripts/tree-traces.js:178-207, as the
reference's test/tree-test.js:18-67 does: extent bits, shape with leaf chains,
and visit / visitAfter records of (leaf head or -1, bound bits)."""

import ctypes
import struct

import numpy as np

NONE = 0xFFFFFFFF

CELL = np.dtype([("child", "<u4", (4,)), ("first", "<u4"), ("count", "<u4"),
                 ("x0", "<f8"), ("y0", "<f8"), ("x1", "<f8"), ("y1", "<f8"),
                 ("x", "<f8"), ("y", "<f8"), ("value", "<f8"), ("r", "<f8")])
assert CELL.itemsize == 88


def bits(v):
    return struct.pack("<d", float(v))[::-1].hex()


def unbits(s):
    return struct.unpack(">d", bytes.fromhex(s))[0]


def read_tree(core, h):
    """(extent, cells ndarray, perm ndarray, m) of the core's last tree."""
    n_cells = core.tree_info(h, 0)
    size = core.tree_info(h, 5)
    assert size == 88
    m = core.tree_info(h, 6)
    ext = (ctypes.c_double * 4).from_address(core.tree_info(h, 3))
    extent = [ext[i] for i in range(4)]
    cells = np.zeros(0, dtype=CELL)
    if n_cells:
        raw = (ctypes.c_char * (n_cells * 88)).from_address(core.tree_info(h, 2))
        cells = np.frombuffer(bytes(raw), dtype=CELL).copy()
    perm = np.zeros(0, dtype=np.uint32)
    if m:
        praw = (ctypes.c_uint32 * m).from_address(core.tree_info(h, 4))
        perm = np.array(praw, dtype=np.uint32)
    return extent, cells, perm, m


def order(core, h, after, n_cells):
    out = (ctypes.c_uint32 * max(n_cells, 1))()
    assert core.tree_order(h, after, ctypes.addressof(out)) == 0
    return [out[i] for i in range(n_cells)]


def core_trace(core, xs, ys):
    """Builds the tree of points (xs, ys) in a fresh simulation and returns
    the trace dict; the simulation is destroyed afterwards."""
    from d3force_amd.core import _view
    n = len(xs)
    h = core.sim_create()
    assert h
    assert core.sim_set_node_count(h, n) >= 0
    if n:
        _view(core.sim_ptr(h, 0), np.float64, n)[:] = xs
        _view(core.sim_ptr(h, 1), np.float64, n)[:] = ys
    assert core.sim_sync(h, 0) >= 0
    assert core.tree_build(h, 0) >= 0
    extent, cells, perm, m = read_tree(core, h)
    n_cells = len(cells)
    visit = order(core, h, 0, n_cells) if n_cells else []
    after = order(core, h, 1, n_cells) if n_cells else []
    root = core.tree_info(h, 1)
    core.sim_destroy(h)

    def is_leaf(c):
        return all(int(k) == NONE for k in cells[c]["child"])

    def chain(c):
        first, count = int(cells[c]["first"]), int(cells[c]["count"])
        return [int(perm[k]) for k in range(first + count - 1, first - 1, -1)]

    def shape(c):
        if c == NONE:
            return None
        if is_leaf(c):
            return {"l": chain(c)}
        return [shape(int(k)) for k in cells[c]["child"]]

    def rec(c):
        cell = cells[c]
        return [chain(c)[0] if is_leaf(c) else -1,
                bits(cell["x0"]), bits(cell["y0"]), bits(cell["x1"]), bits(cell["y1"])]

    return {
        "extent": [bits(v) for v in extent],
        "shape": shape(root) if n_cells else None,
        "visit": [rec(c) for c in visit],
        "after": [rec(c) for c in after],
        "cells": cells,
        "perm": perm,
    }


def load_traces():
    import gzip
    import json
    import os
    import sys
    sys.setrecursionlimit(100000)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tree-traces.json.gz")
    with gzip.open(path, "rt") as f:
        return json.load(f)
