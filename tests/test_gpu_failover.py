"""The failure path of the reference's threaded core (pool.js:73-84,
root.zig:202-209) on the MI355X library, and the facade's copy path around
a force with no native slot (simulation.js:66-77, core.js:121-140).

* A step that returns WORKER_FAILED or OUT_OF_MEMORY names its cause
  (core_last_error) and, with the facade's failover, is rerun on a fresh
  simulation of the same library from the snapshot taken before it, so the
  ticks stay bit-identical to the oracle's. Failures are injected on the host
  (core_fail_next): no device work is done by a failed step, nothing faults.
* sim_snapshot takes vx, vy (and x, y) as they were before a step.
* A Python force (no slot) sees the nodes through copy-out / copy-in.
"""

import numpy as np
import pytest

import d3force_amd as d3
from d3force_amd import core as C
from parity import same
from pins import Raw
from scenarios import random_graph

pytestmark = pytest.mark.gpu


def workload(core, n=3000, seed=3, custom=False):
    src, tgt = random_graph(n, 2 * n, seed=seed)
    nodes = d3.Nodes(n)
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)))
           .force("center", d3.forceCenter()))
    if custom:
        def pull(alpha):  # a force with no native slot: reads and writes the node table
            nodes.vx -= nodes.x * (0.01 * alpha)
            nodes.vy -= np.sign(nodes.y) * alpha
        sim.force("custom", pull)
        sim.force("collide", d3.forceCollide(2.0))
    return sim, nodes


def test_injected_failure_names_its_cause(gpu_core):
    print("gpu failover: status and cause", flush=True)
    s = Raw(gpu_core, 100)
    s.put(np.arange(100.0), np.zeros(100))
    fh = s.force(C.X, 100)
    s.f64(fh, C.F0, 100)[:] = 0.1
    s.f64(fh, C.F1, 100)[:] = 0.0
    assert gpu_core.core_fail_next(C.BAD_ARGUMENT, 0, 1) == C.BAD_ARGUMENT
    assert gpu_core.core_fail_next(C.WORKER_FAILED, 1, 1) == C.OK
    assert s.run(fh, 1.0) == C.OK            # `after` = 1: this step runs
    assert s.run(fh, 1.0) == C.WORKER_FAILED  # the next one fails
    assert b"injected failure" in gpu_core.last_error()
    assert s.run(fh, 1.0) == C.OK            # and only that one
    assert gpu_core.core_fail_next(C.OUT_OF_MEMORY, 0, 1) == C.OK
    assert gpu_core.sim_integrate(s.h, 0.6) == C.OUT_OF_MEMORY
    s.close()
    # Through the facade with failover off, the status raises with its cause.
    sim, nodes = workload(gpu_core, n=500)
    sim.tick()
    ctx = C.context(nodes)
    ctx.failover = False
    assert gpu_core.core_fail_next(C.WORKER_FAILED, 1, 1) == C.OK
    with pytest.raises(C.CoreError) as err:
        sim.tick()
    assert "WORKER_FAILED" in str(err.value) and "injected failure" in str(err.value)


def test_snapshot_holds_the_step_inputs(gpu_core):
    print("gpu failover: snapshot round trip", flush=True)
    n = 2000
    r = np.random.default_rng(4)
    x, y = r.uniform(-300, 300, n), r.uniform(-300, 300, n)
    vx, vy = r.uniform(-1, 1, n), r.uniform(-1, 1, n)
    s = Raw(gpu_core, n)
    s.put(x, y, vx, vy)
    fh = s.force(C.MANY_BODY, n)
    s.f64(fh, C.F0, n)[:] = -30
    s.scalar(fh, C.THETA2, 0.81)
    s.scalar(fh, C.DISTANCE_MIN2, 1)
    s.scalar(fh, C.DISTANCE_MAX2, np.inf)
    for positions in (0, 1):
        before = s.get()
        assert gpu_core.sim_snapshot(s.h, positions) == C.OK
        assert s.run(fh, 0.5) == C.OK
        assert gpu_core.sim_integrate(s.h, 0.6) == C.OK
        after = s.get()
        assert not same(after[2], before[2])
        snap = C._view(gpu_core.sim_ptr(s.h, C.P_SNAP), np.float64, 4 * n)
        assert same(snap[:n], before[2]) and same(snap[n:2 * n], before[3])
        if positions:
            assert same(snap[2 * n:3 * n], before[0]) and same(snap[3 * n:], before[1])
    s.close()


@pytest.mark.parametrize("after", [0, 1, 2, 3, 4, 9])
def test_failover_reruns_the_step_bit_identical(gpu_core, oracle_core, after):
    # Steps per tick: charge, link, center, integrate; `after` picks which
    # step of tick 3 fails (0 charge, 1 link, 2 center, 3 integrate, 4 the
    # next tick's charge, 9 a link two ticks on).
    print("gpu failover: rerun after a failed step %d" % after, flush=True)
    a, na = workload(oracle_core)
    b, nb = workload(gpu_core)
    for t in range(6):
        if t == 2:
            assert gpu_core.core_fail_next(C.WORKER_FAILED if after % 2 == 0 else C.OUT_OF_MEMORY, after, 1) == 0
        a.tick()
        b.tick()
        assert same(na.state(), nb.state()), "tick %d" % (t + 1)
    assert C.context(nb).failovers == 1
    assert a.random_source()() == b.random_source()()


def test_custom_force_copy_path(gpu_core, oracle_core):
    # simulation.js:66-77: a force without a native slot runs on the node
    # table between copy-out and copy-in; the native forces around it keep
    # their bits (collide included).
    print("gpu custom force copy path", flush=True)
    a, na = workload(oracle_core, n=2500, custom=True)
    b, nb = workload(gpu_core, n=2500, custom=True)
    for k in (1, 1, 5, 3):
        a.tick(k)
        b.tick(k)
        assert same(na.state(), nb.state())
    assert a.random_source()() == b.random_source()()
