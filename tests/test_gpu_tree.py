"""GPU exact quadtree build: identical to the reference's d3-quadtree traces
and bit-identical (every 88-byte cell, perm, extent) to the oracle's build."""

import math
import sys

import numpy as np
import pytest

from treeview import core_trace, load_traces, unbits

pytestmark = pytest.mark.gpu
sys.setrecursionlimit(100000)
TRACES = load_traces()


def _same_cells(a, b):
    assert len(a) == len(b), "cell count %d vs %d" % (len(a), len(b))
    if len(a):
        assert a.tobytes() == b.tobytes(), "cells differ at %s" % np.nonzero(a != b)[0][:10]


@pytest.mark.parametrize("name", sorted(TRACES))
def test_gpu_tree_matches_d3_quadtree_trace(gpu_core, name):
    print("gpu tree trace", name, flush=True)
    ref = TRACES[name]
    xs = [unbits(b) for b in ref["x"]]
    ys = [unbits(b) for b in ref["y"]]
    got = core_trace(gpu_core, xs, ys)
    if name in ("empty", "all-nan"):
        assert all(math.isnan(unbits(b)) for b in got["extent"])
    else:
        assert got["extent"] == ref["extent"]
    assert got["shape"] == ref["shape"]
    assert got["visit"] == ref["visit"]
    assert got["after"] == ref["after"]


def _sets():
    rng = np.random.default_rng(5)
    out = {}
    for n in (1000, 10000, 100000):
        out["uniform-%d" % n] = (rng.uniform(-5e3, 5e3, n), rng.uniform(-5e3, 5e3, n))
    n = 20000
    out["clustered-dupes"] = (np.round(rng.normal(0, 30, n), 1), np.round(rng.normal(0, 30, n), 1))
    x = rng.uniform(0, 1, 5000)
    y = rng.uniform(0, 1, 5000)
    x[::7] = np.nan
    y[::11] = np.nan
    out["with-nan"] = (x, y)
    # below 2^-32 of the extent: mixed key groups
    base = np.array([0.5 + k * 1e-12 for k in range(50)] + [1024.0, 0.0])
    out["near-duplicates"] = (np.concatenate([base, base + 3e-13]), np.concatenate([base * 0 + 0.25, base * 0 + 0.25 + 1e-13]))
    xs, ys = [], []
    for c in range(64):
        xs += [c * 100 + k * 2e-11 for k in range(40)]
        ys += [c * 50 + (k % 3) * 1e-12 for k in range(40)]
    out["many-mixed-groups"] = (np.array(xs), np.array(ys))
    # 700 distinct points within 1e-6 of each other beside far points: their
    # keys share the high 32 bits, a run longer than the in-place low-half
    # sort takes, so the build sorts all 64 bits; and a run of 200 that it
    # does sort in place. Index order shuffled so stability matters.
    for size, name in ((700, "long-equal-high-run"), (200, "short-equal-high-run")):
        cx = 5.0 + rng.uniform(0, 1e-6, size)
        cy = -3.0 + rng.uniform(0, 1e-6, size)
        fx = rng.uniform(-1000, 1000, 50)
        fy = rng.uniform(-1000, 1000, 50)
        order = rng.permutation(size + 50)
        out[name] = (np.concatenate([cx, fx])[order], np.concatenate([cy, fy])[order])
    out["signed-zeros"] = (np.array([0.0, -0.0, 1.0, -0.0, 0.0, 2.0]), np.array([-0.0, 0.0, 1.0, 0.0, -0.0, 2.0]))
    out["huge"] = (rng.uniform(-1e300, 1e300, 3000), rng.uniform(-1e300, 1e300, 3000))
    out["tiny"] = (rng.uniform(0, 1e-300, 3000), rng.uniform(0, 1e-300, 3000))
    return out


SETS = _sets()


@pytest.mark.parametrize("name", list(SETS))
def test_gpu_tree_bit_identical_to_oracle(gpu_core, oracle_core, name):
    print("gpu tree vs oracle", name, flush=True)
    xs, ys = SETS[name]
    a = core_trace(oracle_core, xs, ys)
    b = core_trace(gpu_core, xs, ys)
    assert a["extent"] == b["extent"]
    _same_cells(a["cells"], b["cells"])
    assert np.array_equal(a["perm"], b["perm"])
    assert a["visit"] == b["visit"]


@pytest.mark.parametrize("name", ["uniform", "clusters"])
def test_gpu_tree_export_after_many_body_tick(gpu_core, oracle_core, name):
    # After an exact many-body tick the exported cells carry the
    # accumulated charges (x, y, value of many_body.zig:101-135), every
    # 88-byte record bit-identical to the oracle's.
    import d3force_amd as d3
    from treeview import read_tree
    print("tree export after a many-body tick", name, flush=True)
    rng = np.random.default_rng(7)
    if name == "uniform":
        xs, ys = rng.uniform(-300, 300, 5000), rng.uniform(-300, 300, 5000)
    else:
        c = rng.integers(0, 5, 5000)
        xs, ys = c * 100.0 + rng.normal(0, 3, 5000), c * -40.0 + rng.normal(0, 3, 5000)
    out = []
    for core in (oracle_core, gpu_core):
        nodes = d3.Nodes(len(xs), x=xs.copy(), y=ys.copy())
        sim = d3.forceSimulation(nodes, core=core).force("charge", d3.forceManyBody()).stop()
        sim.tick(2)
        out.append(read_tree(core, d3.core.context(nodes).h))
    (ea, ca, pa, ma), (eb, cb, pb, mb) = out
    assert ea == eb and ma == mb
    assert np.array_equal(pa, pb)
    assert np.any(ca["value"] != 0)
    _same_cells(ca, cb)
