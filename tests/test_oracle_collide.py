"""Pins the oracle's exact forceCollide (collide.zig:17-130): the known
answers of the reference's test/d3/collide-test.js:5-55 (d3-force's own
collide tests, compared as that file's assertNodeEqual does, |a - e| < 1e-6,
test/d3/asserts.js:3-15), and bit-for-bit agreement with d3ref's independent
restatement of d3-force collide.js on hand-made layouts (overlapping
clusters, shared coordinates that make a pair's delta exactly zero, NaN and
negative radii, several iterations). The parity corpus scenarios that carry
forceCollide (test/parity/scenarios.js) run in test_oracle_exact.py."""

import math

import numpy as np
import pytest

import d3force_amd as d3
import d3ref
from parity import build_d3ref, build_facade, d3ref_state, same


def near(node, expected, table, i, delta=1e-6):
    got = {"x": table.x[i], "y": table.y[i], "vx": table.vx[i], "vy": table.vy[i]}
    for k, v in expected.items():
        assert abs(got[k] - v) < delta, (i, k, got[k], v)


def test_collide_collides_nodes(oracle_core):  # collide-test.js:5-19
    table = d3.Nodes(3)
    collide = d3.forceCollide(1)
    sim = d3.forceSimulation(table, core=oracle_core).force("collide", collide).stop()
    sim.tick(10)
    near(None, {"x": 7.0710678118654755, "y": 0, "vy": 0, "vx": 0}, table, 0)
    near(None, {"x": -9.03088751750192, "y": 8.27303273571596, "vy": 0, "vx": 0}, table, 1)
    near(None, {"x": 1.3823220809823638, "y": -15.750847141167634, "vy": 0, "vx": 0}, table, 2)
    collide.radius(100)
    sim.tick(10)
    near(None, {"x": 174.08616723117228, "y": 66.51743051995625, "vy": 0.26976816231064354,
                "vx": 0.677346615710878}, table, 0)
    near(None, {"x": -139.73606544743998, "y": 95.69860503079263, "vy": 0.3545632444404687,
                "vx": -0.5300880593105067}, table, 1)
    near(None, {"x": -34.9275994083864, "y": -169.69384995620052, "vy": -0.6243314067511122,
                "vx": -0.1472585564003713}, table, 2)


def test_collide_respects_fixed_positions(oracle_core):  # collide-test.js:22-32
    table = d3.Nodes.from_dicts([{"fx": 0.0, "fy": 0.0}, {}, {}])
    collide = d3.forceCollide(1)
    sim = d3.forceSimulation(table, core=oracle_core).force("collide", collide).stop()
    sim.tick(10)
    near(None, {"x": 0, "y": 0, "vy": 0, "vx": 0}, table, 0)
    collide.radius(100)
    sim.tick(10)
    near(None, {"x": 0, "y": 0, "vy": 0, "vx": 0}, table, 0)


def test_collide_jiggles_equal_positions(oracle_core):  # collide-test.js:34-44
    table = d3.Nodes.from_dicts([{"x": 0.0, "y": 0.0}, {"x": 0.0, "y": 0.0}])
    sim = d3.forceSimulation(table, core=oracle_core).force("collide", d3.forceCollide(1)).stop()
    sim.tick()
    assert table.x[0] != table.x[1]
    assert table.y[0] != table.y[1]
    assert table.vx[0] == -table.vx[1]
    assert table.vy[0] == -table.vy[1]


def test_collide_jiggles_in_a_reproducible_way(oracle_core):  # collide-test.js:46-50
    table = d3.Nodes.from_dicts([{"x": 0.0, "y": 0.0} for _ in range(10)])
    d3.forceSimulation(table, core=oracle_core).force("collide", d3.forceCollide()).stop().tick(50)
    near(None, {"x": -5.371433857229194, "y": -2.6644608278592576, "vy": 0, "vx": 0}, table, 0)


def layouts():
    rng = np.random.default_rng(11)
    out = {}
    pts = rng.normal(0, 6, size=(150, 2))
    out["cluster"] = ([{"x": float(a), "y": float(b)} for a, b in pts], 3.0, 1.0, 1)
    grid = [{"x": float(i % 12) * 2.0, "y": float(i // 12) * 2.0} for i in range(144)]
    out["grid-shared-axes"] = (grid, 1.5, 1.0, 2)
    pts = rng.uniform(-40, 40, size=(200, 2))
    radii = [math.nan if i % 9 == 0 else (-2.0 if i % 13 == 0 else 1 + (i % 5)) for i in range(200)]
    out["nan-and-negative-radii"] = ([{"x": float(a), "y": float(b)} for a, b in pts], radii, 0.7, 3)
    return out


@pytest.mark.parametrize("name", list(layouts()))
def test_collide_matches_d3_restatement(oracle_core, name):
    nodes, radius, strength, iterations = layouts()[name]
    spec = [("collide", {"radius": (lambda i: radius[i]) if isinstance(radius, list) else radius,
                         "strength": strength, "iterations": iterations}), ("center", {})]
    a = build_d3ref(nodes, spec)
    b, table = build_facade(oracle_core, nodes, spec)
    for t in range(25):
        a.tick()
        b.tick()
        assert same(d3ref_state(a), table.state()), "%s tick %d" % (name, t + 1)
    assert a.random() == b.random_source()()
