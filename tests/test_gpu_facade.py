"""GPU facade behaviour a d3 user relies on between ticks, bit-identical to
the oracle: precision() switched mid-run (simulation.js:187-200; fast mode
applies from n >= 2000, root.zig:246-248) and nodes() replaced by a longer,
then a shorter array mid-run (the reference's parity scenario "nodes-resize",
test/parity/scenarios.js; node arrays follow n, sim.zig:59-73)."""

import numpy as np
import pytest

import d3force_amd as d3
from parity import same
from scenarios import random_graph

pytestmark = pytest.mark.gpu


def test_gpu_precision_switch_mid_run(gpu_core, oracle_core):
    # Link and collide are bit-identical in both modes, so the whole run is.
    print("gpu precision switch", flush=True)
    n = 3000
    src, tgt = random_graph(n, 4000, seed=3)
    out = []
    for core in (oracle_core, gpu_core):
        nodes = d3.Nodes(n)
        sim = (d3.forceSimulation(nodes, core=core).force("link", d3.forceLink((src, tgt)))
               .force("collide", d3.forceCollide(4)).force("center", d3.forceCenter()).stop())
        states = []
        for mode in ("exact", "fast", "exact", "fast"):
            sim.precision(mode)
            for _ in range(5):
                sim.tick()
                states.append(nodes.state())
        out.append(states)
    for t, (a, b) in enumerate(zip(*out)):
        assert same(a, b), "tick %d" % (t + 1)


def test_gpu_nodes_resize_mid_run(gpu_core, oracle_core):
    print("gpu nodes resize", flush=True)
    rng = np.random.default_rng(8)
    first = rng.uniform(-250, 250, size=(100, 2))
    extra = rng.uniform(-250, 250, size=(80, 2))
    out = []
    for core in (oracle_core, gpu_core):
        nodes = d3.Nodes(100, x=first[:, 0].copy(), y=first[:, 1].copy())
        sim = (d3.forceSimulation(nodes, core=core).force("charge", d3.forceManyBody())
               .force("collide", d3.forceCollide(5)).force("x", d3.forceX()).stop())
        states = []
        for t in range(200):
            if t == 50:
                cur = sim.nodes()
                nodes = d3.Nodes(180, x=np.concatenate([cur.x, extra[:, 0]]), y=np.concatenate([cur.y, extra[:, 1]]),
                                 vx=np.concatenate([cur.vx, np.zeros(80)]), vy=np.concatenate([cur.vy, np.zeros(80)]))
                sim.nodes(nodes)
            if t == 150:
                cur = sim.nodes()
                nodes = d3.Nodes(150, x=cur.x[30:].copy(), y=cur.y[30:].copy(), vx=cur.vx[30:].copy(),
                                 vy=cur.vy[30:].copy())
                sim.nodes(nodes)
            sim.tick()
            states.append(nodes.state())
        out.append(states)
    for t, (a, b) in enumerate(zip(*out)):
        assert same(a, b), "tick %d" % (t + 1)
