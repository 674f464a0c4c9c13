"""Pins the oracle's exact quadtree build against the reference's own golden
traces of d3-quadtree (test/fixtures/tree-traces.json, recorded by
scripts/tree-traces.js from d3-quadtree 3.0.1) and the Zig unit tests of
src/core/tests.zig:16-116: extent after cover, tree shape with leaf chains,
and the exact visit / visitAfter callback order with bound bits."""

import math
import sys

import numpy as np
import pytest

from treeview import core_trace, load_traces, unbits, NONE

sys.setrecursionlimit(100000)
TRACES = load_traces()


def _compare(name, ref, got):
    # Extents of an empty tree are NaN in both (tree-test.js:75).
    if name in ("empty", "all-nan"):
        assert all(math.isnan(unbits(b)) for b in got["extent"])
    else:
        assert got["extent"] == ref["extent"], name + ": extent"
    assert got["shape"] == ref["shape"], name + ": shape and leaf chains"
    assert got["visit"] == ref["visit"], name + ": visit order and bounds"
    assert got["after"] == ref["after"], name + ": visitAfter order and bounds"


@pytest.mark.parametrize("name", sorted(TRACES))
def test_tree_matches_d3_quadtree_trace(oracle_core, name):
    ref = TRACES[name]
    xs = [unbits(b) for b in ref["x"]]
    ys = [unbits(b) for b in ref["y"]]
    _compare(name, ref, core_trace(oracle_core, xs, ys))


def test_trace_fixture_covers_the_reference_sets():
    assert len(TRACES) == 19


# src/core/tests.zig:16-25
def test_empty_and_all_nan_give_an_empty_tree(oracle_core):
    t = core_trace(oracle_core, [], [])
    assert t["shape"] is None
    t = core_trace(oracle_core, [math.nan, 1.0], [0.0, math.nan])
    assert t["shape"] is None
    assert all(math.isnan(unbits(b)) for b in t["extent"])


# tests.zig:27-34
def test_single_point_extent_floor_plus_one(oracle_core):
    t = core_trace(oracle_core, [3.25], [-7.5])
    assert len(t["cells"]) == 1
    assert [unbits(b) for b in t["extent"]] == [3, -8, 4, -7]


# tests.zig:36-49
def test_coincident_points_form_one_leaf_descending_chain(oracle_core):
    t = core_trace(oracle_core, [1, 1, -0.0, 1], [2, 2, 5, 2])
    found = False
    for c in t["cells"]:
        if all(int(k) == NONE for k in c["child"]) and c["count"] == 3:
            assert list(t["perm"][c["first"]:c["first"] + 3]) == [0, 1, 3]
            found = True
    assert found


# tests.zig:51-60
def test_two_points_split_to_first_separating_level(oracle_core):
    t = core_trace(oracle_core, [0, 1], [0, 1])
    assert [unbits(b) for b in t["extent"]] == [0, 0, 2, 2]
    root = t["cells"][0]
    assert root["child"][0] != NONE and root["child"][3] != NONE
    assert root["child"][1] == NONE and root["child"][2] == NONE


# tests.zig:62-76
def test_near_duplicates_below_key_resolution_separate(oracle_core):
    t = core_trace(oracle_core, [0, 1024, 0.5, 0.5 + 1e-12], [0, 1024, 0.25, 0.25])
    leaves = [c for c in t["cells"] if all(int(k) == NONE for k in c["child"])]
    assert len(leaves) == 4
    assert min(c["x1"] - c["x0"] for c in leaves) < 2048.0 / 2 ** 32


# tests.zig:78-95 (sorted and stable), with the oracle's sorted keys.
def test_sorted_keys_are_sorted_and_stable(oracle_core):
    import ctypes
    from d3force_amd.core import _view
    rng = np.random.default_rng(42)
    xs = rng.integers(-20, 20, 500).astype(np.float64)
    ys = np.where(np.arange(500) % 2 == 0, 3.0, rng.integers(-20, 20, 500).astype(np.float64))
    core = oracle_core
    h = core.sim_create()
    core.sim_set_node_count(h, 500)
    _view(core.sim_ptr(h, 0), np.float64, 500)[:] = xs
    _view(core.sim_ptr(h, 1), np.float64, 500)[:] = ys
    core.tree_build(h, 0)
    m = core.tree_info(h, 6)
    keys = np.array((ctypes.c_uint64 * m).from_address(core.tree_keys(h)))
    perm = np.array((ctypes.c_uint32 * m).from_address(core.tree_info(h, 4)))
    core.sim_destroy(h)
    assert np.all(keys[:-1] <= keys[1:])
    eq = keys[:-1] == keys[1:]
    assert np.all(perm[:-1][eq] < perm[1:][eq])


# tests.zig:97-116
def test_visit_is_preorder_and_visit_after_is_postorder(oracle_core):
    from treeview import order
    from d3force_amd.core import _view
    core = oracle_core
    xs, ys = [0, 3, 1, 2, 3.5], [0, 3, 2, 1, 0.5]
    h = core.sim_create()
    core.sim_set_node_count(h, 5)
    _view(core.sim_ptr(h, 0), np.float64, 5)[:] = xs
    _view(core.sim_ptr(h, 1), np.float64, 5)[:] = ys
    core.tree_build(h, 0)
    n = core.tree_info(h, 0)
    assert order(core, h, 0, n) == list(range(n))
    after = order(core, h, 1, n)
    assert after[-1] == core.tree_info(h, 1)
    core.sim_destroy(h)


def test_handles(oracle_core):
    # tree-test.js:93-106: destroyed and unknown handles are rejected.
    core = oracle_core
    h = core.sim_create()
    assert core.sim_destroy(h) == 0
    assert core.sim_destroy(h) == -1
    assert core.tree_build(12345, 0) == -1
