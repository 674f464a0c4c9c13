"""Whole ticks replayed from a captured HIP graph (sim_ticks, the facade's
default for tick(k) when every force runs in the core) against the per-force
loop of the same library, simulation.js:63-82 (each force in order, then
integrate): x, y, vx, vy, alpha and the random source's position identical
after every call. The replays that must be rolled back and run eagerly are
forced from the data: two pinned nodes on one point give a jiggle in the
exact traversal and a zero-length link on every tick (many_body.zig:77-83,
link.zig:23-27). The per-force loop itself is checked against the oracle by
the other GPU tests; the 20k-node exact case here is also checked against it."""

import numpy as np
import pytest

import d3force_amd as d3
from d3force_amd.core import context
from parity import same
from scenarios import random_graph

pytestmark = pytest.mark.gpu


def _sim(core, n, m, captured, seed=1, precision=None, pinned_pair=False, extra=()):
    src, tgt = random_graph(n, m, seed=seed)
    nodes = d3.Nodes(n)
    if pinned_pair:
        # nodes 0 and 1 pinned on one point, with a link between them
        nodes.fix(fx=np.array([5.0, 5.0]), fy=np.array([-3.0, -3.0]), where=np.array([0, 1]))
        src = np.concatenate([[0], src])
        tgt = np.concatenate([[1], tgt])
    sim = (d3.forceSimulation(nodes, core=core).force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt))).force("center", d3.forceCenter()))
    for name, f in extra:
        sim.force(name, f)
    if precision:
        sim.precision(*precision)
    if core.sim_ticks is not None:
        sim.captured_ticks(captured)
    return sim, nodes


def _runs(cores, calls, **kw):
    """States after each tick(k) call for k in calls, one list per (core, captured)."""
    out = []
    for core, captured in cores:
        sim, nodes = _sim(core, captured=captured, **kw)
        states = []
        for k in calls:
            sim.tick(k)
            states.append((nodes.state(), sim.random_source().state[0], sim.alpha()))
        counters = context(nodes).tick_counters() if core.sim_ticks is not None else None
        out.append((states, counters))
    return out


def _same_runs(a, b, what):
    for c, (sa, sb) in enumerate(zip(a, b)):
        assert same(sa[0], sb[0]), "%s: state after call %d" % (what, c + 1)
        assert sa[1] == sb[1], "%s: random source after call %d" % (what, c + 1)
        assert sa[2] == sb[2], "%s: alpha after call %d" % (what, c + 1)


def test_captured_exact_ticks_match_per_force_loop_and_oracle(gpu_core, oracle_core):
    print("captured exact 20k", flush=True)
    calls = [1, 1, 7, 12]
    (eager, _), (cap, counters), (orc, _) = _runs(
        [(gpu_core, False), (gpu_core, True), (oracle_core, False)], calls, n=20000, m=40000)
    _same_runs(eager, cap, "captured vs per-force")
    _same_runs(orc, cap, "captured vs oracle")
    captures, replayed, redone = counters
    assert captures >= 1 and replayed >= 18 and redone == 0, counters


def test_captured_fast_ticks_match_per_force_loop(gpu_core):
    # fast precision draws a seed per manyBody and link run (core.js:213-231)
    print("captured fast 50k", flush=True)
    calls = [1, 5, 10]
    (eager, _), (cap, counters) = _runs([(gpu_core, False), (gpu_core, True)], calls, n=50000, m=100000,
                                        precision=("fast", {"order": 8}))
    _same_runs(eager, cap, "captured vs per-force (fast)")
    assert counters[1] >= 14, counters


def test_captured_ticks_roll_back_jiggles_and_zero_links(gpu_core, oracle_core):
    print("captured exact, coincident pinned pair", flush=True)
    calls = [1, 6, 6]
    (eager, _), (cap, counters), (orc, _) = _runs(
        [(gpu_core, False), (gpu_core, True), (oracle_core, False)], calls, n=3000, m=5000, pinned_pair=True)
    _same_runs(eager, cap, "captured vs per-force (jiggles)")
    _same_runs(orc, cap, "captured vs oracle (jiggles)")
    assert counters[2] >= 1, counters


def test_captured_fast_ticks_roll_back_with_seeds_in_order(gpu_core):
    # below 2000 nodes fast precision runs the exact kernels but still draws
    # a seed per run: the jiggles of a rolled-back tick come between them
    print("captured fast-precision small graph with jiggles", flush=True)
    calls = [1, 4, 4]
    (eager, _), (cap, counters) = _runs([(gpu_core, False), (gpu_core, True)], calls, n=1500, m=2500,
                                        pinned_pair=True, precision=("fast", {}))
    _same_runs(eager, cap, "captured vs per-force (fast precision, exact kernels)")
    assert counters[2] >= 1, counters


def test_captured_ticks_with_axis_and_radial_forces(gpu_core):
    print("captured exact with x, y, radial", flush=True)
    extra = [("x", d3.forceX(10)), ("y", d3.forceY(-20)), ("radial", d3.forceRadial(100))]
    calls = [1, 5, 5]
    (eager, _), (cap, counters) = _runs([(gpu_core, False), (gpu_core, True)], calls, n=5000, m=8000, extra=extra)
    _same_runs(eager, cap, "captured vs per-force (x, y, radial)")
    assert counters[1] >= 8, counters


def test_captured_ticks_follow_a_changed_parameter(gpu_core):
    # a scalar changed between calls: the tick is captured again
    print("captured exact, parameter change", flush=True)
    out = []
    for captured in (False, True):
        sim, nodes = _sim(gpu_core, 4000, 6000, captured)
        states = []
        for k, theta in ((3, 0.9), (4, 0.5), (4, 1.2)):
            sim.force("charge").theta(theta)
            sim.tick(k)
            states.append((nodes.state(), sim.random_source().state[0], sim.alpha()))
        out.append((states, context(nodes).tick_counters()))
    _same_runs(out[0][0], out[1][0], "captured vs per-force (theta changed)")
    assert out[1][1][0] >= 3, out[1][1]
