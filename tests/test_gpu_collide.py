"""GPU exact forceCollide (collide.zig:17-130): bit-identical (Object.is on x,
y, vx, vy after every tick, and the random stream's position) to the oracle.
Covers d3's own collide tests (test/d3/collide-test.js:5-55), layouts with
shared coordinates (zero deltas that draw jiggles), NaN and negative radii,
several iterations, and layouts of 20,000 and 200,000 nodes whose overlaps the
parallel pass resolves without the serial kernel. The parity corpus scenarios
that carry forceCollide run in test_gpu_exact.py."""

import numpy as np
import pytest

import d3force_amd as d3
from parity import build_facade, same
from scenarios import random_graph
from test_oracle_collide import layouts

pytestmark = pytest.mark.gpu


def run_both(gpu_core, oracle_core, make, ticks):
    a, ta = make(oracle_core)
    b, tb = make(gpu_core)
    for t in range(ticks):
        a.tick()
        b.tick()
        assert same(ta.state(), tb.state()), "tick %d" % (t + 1)
    assert a.random_source()() == b.random_source()(), "random stream position"
    return b


def test_gpu_collide_d3_collides_nodes(gpu_core, oracle_core):  # collide-test.js:5-19
    print("gpu collide d3 test 1", flush=True)
    ta, tb = d3.Nodes(3), d3.Nodes(3)
    ca, cb = d3.forceCollide(1), d3.forceCollide(1)
    a = d3.forceSimulation(ta, core=oracle_core).force("collide", ca).stop()
    b = d3.forceSimulation(tb, core=gpu_core).force("collide", cb).stop()
    a.tick(10)
    b.tick(10)
    assert same(ta.state(), tb.state())
    assert abs(tb.x[1] - -9.03088751750192) < 1e-6 and abs(tb.y[2] - -15.750847141167634) < 1e-6
    ca.radius(100)
    cb.radius(100)
    a.tick(10)
    b.tick(10)
    assert same(ta.state(), tb.state())
    assert abs(tb.x[0] - 174.08616723117228) < 1e-6 and abs(tb.vx[2] - -0.1472585564003713) < 1e-6


def test_gpu_collide_d3_jiggles(gpu_core, oracle_core):  # collide-test.js:34-50
    print("gpu collide d3 jiggle tests", flush=True)

    def two(core):
        t = d3.Nodes.from_dicts([{"x": 0.0, "y": 0.0}, {"x": 0.0, "y": 0.0}])
        return d3.forceSimulation(t, core=core).force("collide", d3.forceCollide(1)).stop(), t

    run_both(gpu_core, oracle_core, two, 1)

    def ten(core):
        t = d3.Nodes.from_dicts([{"x": 0.0, "y": 0.0} for _ in range(10)])
        return d3.forceSimulation(t, core=core).force("collide", d3.forceCollide()).stop(), t

    run_both(gpu_core, oracle_core, ten, 50)


@pytest.mark.parametrize("name", list(layouts()))
def test_gpu_collide_layouts(gpu_core, oracle_core, name):
    print("gpu collide layout", name, flush=True)
    nodes, radius, strength, iterations = layouts()[name]
    spec = [("collide", {"radius": (lambda i: radius[i]) if isinstance(radius, list) else radius,
                         "strength": strength, "iterations": iterations}), ("center", {})]
    run_both(gpu_core, oracle_core, lambda core: build_facade(core, nodes, spec), 25)


def scatter(n, seed, spread):
    rng = np.random.default_rng(seed)
    xy = rng.uniform(-spread, spread, size=(n, 2))
    radii = rng.uniform(0.5, 1.5, size=n)
    return xy, radii


@pytest.mark.parametrize("n,spread,ticks", [(20000, 300.0, 8), (200000, 1000.0, 3), (1000000, 2120.6, 3)])
def test_gpu_collide_parallel_pass(gpu_core, oracle_core, n, spread, ticks):
    # Uniform points with radii U[0.5, 1.5]: most nodes overlap a neighbour,
    # by less than a radius, so every iteration runs as the dataflow (no
    # serial pass) and matches the oracle bit for bit. The 1M case is the
    # benchmark's collide workload (0.0556 nodes per unit area).
    print("gpu collide parallel", n, flush=True)
    xy, radii = scatter(n, 5, spread)

    def make(core):
        t = d3.Nodes(n, x=xy[:, 0].copy(), y=xy[:, 1].copy())
        sim = (d3.forceSimulation(t, core=core)
               .force("collide", d3.forceCollide(radii).strength(0.7).iterations(2)).stop())
        return sim, t

    b = run_both(gpu_core, oracle_core, make, ticks)
    st = gpu_core.stats(d3.core.context(b.nodes()).h)
    assert st["collide_serial"] == 0, "stop reasons %d" % st["collide_status"]


def test_gpu_collide_jiggles_continue_the_parallel_pass(gpu_core, oracle_core):
    # 20,000 uniform nodes with 20 pairs placed on the same point: those
    # pairs draw jiggles (collide.zig:62-69), so their nodes stop; the rest of
    # the pass finishes in parallel and the unfinished nodes run serially in
    # index order, drawing from the host's random as the reference does.
    print("gpu collide jiggle continuation", flush=True)
    n = 20000
    xy, radii = scatter(n, 9, 300.0)
    rng = np.random.default_rng(4)
    src = rng.choice(n, 40, replace=False)
    xy[src[20:]] = xy[src[:20]]

    def make(core):
        t = d3.Nodes(n, x=xy[:, 0].copy(), y=xy[:, 1].copy())
        sim = d3.forceSimulation(t, core=core).force("collide", d3.forceCollide(radii).strength(0.7)).stop()
        return sim, t

    b = run_both(gpu_core, oracle_core, make, 4)
    st = gpu_core.stats(d3.core.context(b.nodes()).h)
    assert st["collide_serial"] == 0, "stop reasons %d" % st["collide_status"]
    assert st["collide_status"] & 4, "no jiggle was drawn"


EDGE = {
    "no-nodes": (0, None, 1.0, 1.0, 1),
    "one-node": (1, None, 3.0, 1.0, 1),
    "two-coincident": (2, "same", 3.0, 1.0, 2),
    "all-nan-positions": (20, "nan", 3.0, 1.0, 1),
    "zero-strength": (300, None, 8.0, 0.0, 1),
    "negative-strength": (300, None, 8.0, -0.5, 1),
    "zero-iterations": (300, None, 8.0, 1.0, 0),
    "infinite-radius-one-node": (300, "inf", 4.0, 1.0, 1),
}


@pytest.mark.parametrize("name", list(EDGE))
def test_gpu_collide_edge_cases(gpu_core, oracle_core, name):
    print("gpu collide edge", name, flush=True)
    n, kind, radius, strength, iterations = EDGE[name]
    rng = np.random.default_rng(12)
    xy = rng.uniform(-60, 60, size=(n, 2))
    if kind == "same":
        xy[:] = 1.5
    if kind == "nan":
        xy[:] = np.nan
    radii = np.full(n, radius)
    if kind == "inf":
        radii[7] = np.inf

    def make(core):
        t = d3.Nodes(n, x=xy[:, 0].copy(), y=xy[:, 1].copy()) if n else d3.Nodes(0)
        sim = (d3.forceSimulation(t, core=core)
               .force("collide", d3.forceCollide(radii).strength(strength).iterations(iterations)).stop())
        return sim, t

    run_both(gpu_core, oracle_core, make, 3)


def test_gpu_collide_in_batched_ticks(gpu_core, oracle_core):
    # The reference's parity scenario "batched-ticks" (test/parity/scenarios.js):
    # tick(7) batches, one copy in and out around 7 ticks, with forceCollide
    # after the default forces on 200 nodes placed by the phyllotaxis.
    print("gpu collide batched ticks", flush=True)
    src, tgt = random_graph(200, 300, seed=21)
    out = []
    for core in (oracle_core, gpu_core):
        nodes = d3.Nodes(200)
        sim = (d3.forceSimulation(nodes, core=core).force("charge", d3.forceManyBody())
               .force("link", d3.forceLink((src, tgt))).force("center", d3.forceCenter())
               .force("collide", d3.forceCollide(3)).stop())
        states = []
        for _ in range(6):
            sim.tick(7)
            states.append(nodes.state())
        out.append((states, sim.random_source()()))
    for k, (a, b) in enumerate(zip(out[0][0], out[1][0])):
        assert same(a, b), "batch %d" % (k + 1)
    assert out[0][1] == out[1][1]
