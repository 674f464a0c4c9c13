"""Pins the oracle's FP64 fast mode (FMM many-body, adjacency-sweep link,
uniform-grid collide, center, integrate, and the facade's per-run seeds)
against the reference's golden SHA-256 digests of x, y, vx, vy
(test/fixtures/fast-golden.json, checked by test/threads/threads-test.js:107-121
over the layouts of test/threads/scenarios.js)."""

import json
import os

import pytest

import d3force_amd as d3
from scenarios import CHECKPOINTS, FAST_LAYOUTS, digest, fast_scenario

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fast-golden.json")))


def run_fast(core, name, last):
    xs, ys, src, tgt = fast_scenario(name)
    nodes = d3.Nodes(len(xs), x=xs, y=ys)
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)))
           .force("collide", d3.forceCollide(3))
           .force("center", d3.forceCenter())
           .stop())
    sim.precision("fast")
    out = {}
    for t in range(1, last + 1):
        sim.tick()
        if t in CHECKPOINTS:
            out[t] = digest(nodes)
    return out


@pytest.mark.parametrize("name", list(FAST_LAYOUTS))
def test_fast_mode_matches_reference_digests_to_tick_10(oracle_core, name):
    got = run_fast(oracle_core, name, 10)
    for t in (1, 10):
        assert got[t] == GOLDEN[name][str(t)], "%s tick %d" % (name, t)


@pytest.mark.slow
@pytest.mark.parametrize("name", list(FAST_LAYOUTS))
def test_fast_mode_matches_reference_digests_to_tick_300(oracle_core, name):
    got = run_fast(oracle_core, name, 300)
    for t in CHECKPOINTS:
        assert got[t] == GOLDEN[name][str(t)], "%s tick %d" % (name, t)
