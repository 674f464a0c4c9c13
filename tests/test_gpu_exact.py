"""GPU exact mode: bit-identical (Object.is on x, y, vx, vy after every tick,
and the random stream's position) to the oracle, over the reference's parity
corpus scenarios and the benchmark workloads."""

import numpy as np
import pytest

import d3force_amd as d3
from parity import SCENARIOS, build_facade, same
from scenarios import generator, pinned, random_graph

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(SCENARIOS))
def test_gpu_exact_tick_matches_oracle(gpu_core, oracle_core, name):
    print("gpu exact scenario", name, flush=True)
    for seed in (1, 2):
        nodes, forces = SCENARIOS[name](generator(seed))
        a, ta = build_facade(oracle_core, nodes, forces)
        b, tb = build_facade(gpu_core, nodes, forces)
        for t in range(40):
            a.tick()
            b.tick()
            assert same(ta.state(), tb.state()), "%s seed %d tick %d" % (name, seed, t + 1)
        assert a.random_source()() == b.random_source()(), "random stream position"


def workload(core, n, m, pin_fraction=0.0, seed=1):
    """The benchmark graph: phyllotaxis start, m links over distinct pairs,
    forceManyBody + forceLink + forceCenter with d3's defaults."""
    src, tgt = random_graph(n, m, seed=seed)
    nodes = d3.Nodes(n)
    if pin_fraction:
        idx, fx, fy = pinned(n, pin_fraction)
        nodes.fix(fx=fx, fy=fy, where=idx)
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)))
           .force("center", d3.forceCenter()))
    return sim, nodes


def test_gpu_exact_10k_300_ticks_bit_identical(gpu_core, oracle_core):
    print("gpu exact 10k x 300", flush=True)
    a, na = workload(oracle_core, 10000, 20000)
    b, nb = workload(gpu_core, 10000, 20000)
    for t in range(300):
        a.tick()
        b.tick()
        assert same(na.state(), nb.state()), "tick %d" % (t + 1)
    assert a.random_source()() == b.random_source()()


def test_gpu_exact_tick_batches_match(gpu_core, oracle_core):
    # tick(k): one copy in and out around k iterations (simulation.js:49-89).
    a, na = workload(oracle_core, 3000, 6000, pin_fraction=0.01)
    b, nb = workload(gpu_core, 3000, 6000, pin_fraction=0.01)
    for k in (1, 7, 20):
        a.tick(k)
        b.tick(k)
        assert same(na.state(), nb.state()), "after tick(%d)" % k


@pytest.mark.timeout(1200)
def test_gpu_exact_1m_pinned_300_ticks_bit_identical(gpu_core, oracle_core):
    # The 1M-node configuration's criterion: x, y, vx, vy identical (Object.is)
    # to the reference after every one of 300 ticks (about 0.3 s per tick for
    # the oracle on 16 threads).
    print("gpu exact 1M pinned x 300", flush=True)
    oracle_core.set_threads(16)
    a, na = workload(oracle_core, 1000000, 2000000, pin_fraction=0.01)
    b, nb = workload(gpu_core, 1000000, 2000000, pin_fraction=0.01)
    for t in range(300):
        a.tick()
        b.tick()
        assert same(na.state(), nb.state()), "tick %d" % (t + 1)
        if (t + 1) % 25 == 0:
            print("  tick %d identical" % (t + 1), flush=True)
    oracle_core.set_threads(1)


def test_gpu_exact_1m_pinned_bit_identical(gpu_core, oracle_core):
    print("gpu exact 1M pinned", flush=True)
    oracle_core.set_threads(16)
    a, na = workload(oracle_core, 1000000, 2000000, pin_fraction=0.01)
    b, nb = workload(gpu_core, 1000000, 2000000, pin_fraction=0.01)
    for t in range(2):
        a.tick()
        b.tick()
        assert same(na.state(), nb.state()), "tick %d" % (t + 1)
    oracle_core.set_threads(1)


def test_gpu_exact_hub_links_many_levels(gpu_core, oracle_core):
    # A hub linked to 700 nodes puts each of its links in a level of its own
    # (link.zig:14-38 runs the links in order): hundreds of small levels,
    # which run in one workgroup per 62 levels, with ordinary levels between
    print("gpu exact hub links", flush=True)
    n = 4000
    src, tgt = random_graph(n, 6000, seed=3)
    hub = np.arange(1, 701, dtype=src.dtype)
    src = np.concatenate([src[:3000], np.zeros(700, src.dtype), src[3000:], hub[::-1]])
    tgt = np.concatenate([tgt[:3000], hub, tgt[3000:], np.full(700, 2, tgt.dtype)])
    states = []
    for core in (oracle_core, gpu_core):
        nodes = d3.Nodes(n)
        sim = (d3.forceSimulation(nodes, core=core).force("charge", d3.forceManyBody())
               .force("link", d3.forceLink((src, tgt)).iterations(2)).force("center", d3.forceCenter()))
        out = []
        for k in (1, 1, 5, 5):
            sim.tick(k)
            out.append(nodes.state())
        states.append(out)
    for c, (a, b) in enumerate(zip(*states)):
        assert same(a, b), "after call %d" % (c + 1)
