"""Batched fast mode: many graphs ticked together (d3force_amd.batch).

- A graph's result in a batch is bit-identical to a batch of that graph alone
  (nothing in a graph's arithmetic depends on the other graphs).
- A graph below 2000 nodes runs the exact kernels, as the reference does
  (root.zig:246-248): bit-identical to the oracle's exact tick of that graph
  alone through the facade in fast precision, random stream position
  included (the facade draws the manyBody and link seeds, core.js:218-222).
- A graph of at least 2000 nodes (fast mode) is bit-identical to the
  single-simulation GPU fast path with manyBody, link and center and d3's
  default random source (the same FMM, link sweep, center sums, integrate and
  seeds).
- The many-body force of every fast graph is within FP32 rounding of the
  FMM's truncation error from an FP64 direct sum with d3's pair term.
- The 4,096-graph workload (n log-uniform in [50, 20000]) passes the same
  checks on a sample of its graphs.
"""

import math

import numpy as np
import pytest

import d3force_amd as d3
from d3force_amd import core as C
from d3force_amd.batch import (Batch, BS_CENTER_STRENGTH, BS_LINK_ITERATIONS, BS_VELOCITY_DECAY, BS_ALPHA,
                               phyllotaxis_local)
from parity import same
from scenarios import sbm_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def graphs():
    # Sizes across the range of the 4096-graph workload, including tiny ones.
    gs = sbm_batch(40, seed=5, n_min=50, n_max=12000)
    return gs


def test_batch_graph_equals_its_batch_of_one(gpu_core, graphs):
    print("batch of", len(graphs), flush=True)
    b = Batch(graphs, core=gpu_core)
    b.tick(6)
    full = b.state()
    for g in (0, 3, 7, len(graphs) - 1):
        one = Batch([graphs[g]], core=gpu_core)
        one.tick(6)
        assert same(b.graph_state(g, full), one.state()), "graph %d (n = %d)" % (g, graphs[g][0])
        one.close()
    b.close()


def test_batch_graph_equals_the_fast_simulation(gpu_core, graphs):
    big = [g for g in range(len(graphs)) if graphs[g][0] >= 2000][:3]
    assert big
    b = Batch(graphs, core=gpu_core)
    b.tick(5)
    full = b.state()
    for g in big:
        n, src, tgt = graphs[g]
        x, y = phyllotaxis_local(n)
        nodes = d3.Nodes(n, x=x, y=y)
        sim = (d3.forceSimulation(nodes, core=gpu_core)
               .force("charge", d3.forceManyBody())
               .force("link", d3.forceLink((src, tgt)))
               .force("center", d3.forceCenter()).stop())
        sim.precision("fast", {"order": 8, "opening": 1.5})
        sim.tick(5)
        assert same(b.graph_state(g, full), nodes.state()), "graph %d (n = %d)" % (g, n)
    b.close()


def direct(x, y, alpha, sample, strength=-30.0):
    fx = np.empty(len(sample))
    fy = np.empty(len(sample))
    for k, i in enumerate(sample):
        dx, dy = x - x[i], y - y[i]
        l = dx * dx + dy * dy
        keep = l > 0
        dx, dy, l = dx[keep], dy[keep], l[keep]
        l = np.where(l < 1, np.sqrt(l), l)
        w = strength * alpha / l
        fx[k], fy[k] = np.sum(dx * w), np.sum(dy * w)
    return fx, fy


def test_batch_many_body_error_vs_direct_sum(gpu_core, graphs):
    # One manyBody-only tick (no link, no centering, velocity kept): v is the
    # force on the starting layout.
    b = Batch(graphs, core=gpu_core)
    b.scalar(BS_LINK_ITERATIONS, 0).scalar(BS_CENTER_STRENGTH, 0).scalar(BS_VELOCITY_DECAY, 1)
    start = b.state()
    b.tick(1)
    alpha = b.scalar(BS_ALPHA)
    st = b.state()
    rng = np.random.default_rng(1)
    worst = 0.0
    for g in range(0, len(graphs), 4):
        if graphs[g][0] < 2000:
            continue  # exact Barnes-Hut (tested against the oracle below)
        s0 = b.graph_state(g, start)
        s1 = b.graph_state(g, st)
        n = s0.shape[1]
        sample = rng.choice(n, min(n, 300), replace=False)
        fx, fy = direct(s0[0], s0[1], alpha, sample)
        e = math.sqrt(np.sum((s1[2][sample] - fx) ** 2 + (s1[3][sample] - fy) ** 2) / np.sum(fx * fx + fy * fy))
        worst = max(worst, e)
    print("  worst per-graph rms vs FP64 direct sum %.2e" % worst, flush=True)
    assert worst < 1e-5
    b.close()


def facade_run(core, n, src, tgt, ticks, x, y, per_tick=None, iterations=1):
    """One graph alone through the facade in fast precision (exact kernels
    below 2000 nodes), from the batch's start (positions x, y)."""
    nodes = d3.Nodes(n, x=x.copy(), y=y.copy())
    sim = (d3.forceSimulation(nodes, core=core)
           .force("charge", d3.forceManyBody())
           .force("link", d3.forceLink((src, tgt)).iterations(iterations))
           .force("center", d3.forceCenter()).stop())
    sim.precision("fast", {"order": 8, "opening": 1.5})
    for t in range(ticks):
        sim.tick()
        if per_tick:
            per_tick(t, nodes.state())
    return sim, nodes


def check_small_graphs_exact(core, oracle, b, graphs, pick, ticks):
    """Graphs `pick` (below 2000 nodes) of batch b, after `ticks` more ticks,
    against the oracle's facade run: x, y, vx, vy after every tick and the
    random stream's position."""
    start = b.state()
    states = []
    for t in range(ticks):
        b.tick(1)
        states.append(b.state())
    lcg = b.random_states()
    for g in pick:
        n, src, tgt = graphs[g]
        assert n < 2000
        s0 = b.graph_state(g, start)

        def each(t, st, g=g):
            assert same(b.graph_state(g, states[t]), st), "graph %d (n = %d) tick %d" % (g, n, t + 1)
        sim, _ = facade_run(oracle, n, src, tgt, ticks, s0[0], s0[1], each)
        assert lcg[g] == sim.random_source().state[0], "graph %d random stream" % g


def test_batch_small_graphs_equal_the_oracle(gpu_core, oracle_core, graphs):
    print("batch small graphs vs oracle", flush=True)
    b = Batch(graphs, core=gpu_core)
    small = [g for g in range(len(graphs)) if graphs[g][0] < 2000]
    assert len(small) >= 10
    check_small_graphs_exact(gpu_core, oracle_core, b, graphs, small, 12)
    b.close()


def test_batch_small_graph_jiggles_equal_the_oracle(gpu_core, oracle_core):
    # Coincident nodes: Barnes-Hut jiggles (replayed in node order from the
    # graph's random source) and a zero-length link (jiggled in link order).
    print("batch small graphs with jiggles", flush=True)
    r = np.random.default_rng(9)
    gs, pos = [], []
    for n in (40, 300, 1500):
        src = r.integers(0, n, 2 * n)
        tgt = (src + 1 + r.integers(0, n - 1, 2 * n)) % n
        x = r.uniform(-100, 100, n)
        y = r.uniform(-100, 100, n)
        x[1::7], y[1::7] = x[0::7][:len(x[1::7])], y[0::7][:len(y[1::7])]
        x[5] = x[6]  # an axis-aligned pair: one zero delta
        src[0], tgt[0] = 1, 0  # a zero-length link
        gs.append((n, src, tgt))
        pos.append((x, y))
    b = Batch(gs, core=gpu_core, positions=pos)
    start = b.state()
    for t in range(6):
        b.tick(1)
    lcg = b.random_states()
    st = b.state()
    for g, (n, src, tgt) in enumerate(gs):
        s0 = b.graph_state(g, start)
        sim, nodes = facade_run(oracle_core, n, src, tgt, 6, s0[0], s0[1])
        assert same(b.graph_state(g, st), nodes.state()), "graph %d" % g
        assert lcg[g] == sim.random_source().state[0]
    b.close()


def test_batch_small_graph_deep_cells_and_link_passes(gpu_core, oracle_core):
    # Points 1e-11 apart (they share all 32 key levels without coinciding, so
    # the tree splits them below the keys), a star and a repeated link (many
    # link levels, one link a level), and three link passes a tick.
    print("batch small graphs: deep cells, link levels, 3 link passes", flush=True)
    r = np.random.default_rng(11)
    gs, pos = [], []
    for n, kind in ((120, "deep"), (900, "deep"), (1999, "star")):
        src = r.integers(0, n, 3 * n)
        tgt = (src + 1 + r.integers(0, n - 1, 3 * n)) % n
        x = r.uniform(-100, 100, n)
        y = r.uniform(-100, 100, n)
        if kind == "deep":  # two pairs: each a one-child chain to about depth 45
            x[[1, 8]] = x[[0, 7]] + 1e-11
            y[[1, 8]] = y[[0, 7]] - 1e-11
        else:
            src[:n - 1], tgt[:n - 1] = 0, np.arange(1, n)
            src[n:n + 300], tgt[n:n + 300] = 5, 6
        gs.append((n, src, tgt))
        pos.append((x, y))
    b = Batch(gs, core=gpu_core, positions=pos)
    b.scalar(BS_LINK_ITERATIONS, 3)
    start = b.state()
    states = []
    for t in range(4):
        b.tick(1)
        states.append(b.state())
    lcg = b.random_states()
    for g, (n, src, tgt) in enumerate(gs):
        s0 = b.graph_state(g, start)

        def each(t, st, g=g):
            assert same(b.graph_state(g, states[t]), st), "graph %d tick %d" % (g, t + 1)
        sim, _ = facade_run(oracle_core, n, src, tgt, 4, s0[0], s0[1], each, iterations=3)
        assert lcg[g] == sim.random_source().state[0]
    b.close()
    # Thirteen such pairs in 120 nodes outgrow the 3 n + 128 cells the kernel
    # holds per graph: the tick fails with its cause, nothing is written.
    n, src, tgt = gs[0]
    x, y = r.uniform(-100, 100, n), r.uniform(-100, 100, n)
    x[1::9], y[1::9] = x[0::9][:len(x[1::9])] + 1e-11, y[0::9][:len(y[1::9])] - 1e-11
    b = Batch([gs[0]], core=gpu_core, positions=[(x, y)])
    before = b.state()
    with pytest.raises(C.CoreError) as err:
        b.tick(1)
    assert "WORKER_FAILED" in str(err.value) and "3 n + 128 cells" in str(err.value)
    assert same(b.state(), before)
    b.close()


@pytest.mark.timeout(600)
def test_batch_4096_graph_workload(gpu_core, oracle_core):
    # The batched configuration: 4,096 SBM graphs, n log-uniform in
    # [50, 20000]; a few ticks, then on a sample of graphs: identity with a
    # batch of one, with the oracle (small graphs, exact) or with the
    # single-simulation GPU fast path (large graphs), and the fast graphs'
    # many-body error against the FP64 direct sum.
    print("batch 4096 graphs", flush=True)
    gs = sbm_batch(4096, seed=1)
    b = Batch(gs, core=gpu_core)
    sizes = np.array([g[0] for g in gs])
    small = np.nonzero(sizes < 2000)[0]
    large = np.nonzero(sizes >= 2000)[0]
    print("  %d graphs below 2000 nodes, %d at or above" % (len(small), len(large)), flush=True)
    start = b.state()
    b.tick(3)
    st = b.state()
    lcg = b.random_states()
    rng = np.random.default_rng(4)
    for g in sorted(rng.choice(len(gs), 6, replace=False).tolist()):
        one = Batch([gs[g]], core=gpu_core, positions=[tuple(b.graph_state(g, start)[:2])])
        one.tick(3)
        assert same(b.graph_state(g, st), one.state()), "graph %d alone" % g
        one.close()
    for g in sorted(rng.choice(small, 8, replace=False).tolist()):
        n, src, tgt = gs[g]
        s0 = b.graph_state(g, start)
        sim, nodes = facade_run(oracle_core, n, src, tgt, 3, s0[0], s0[1])
        assert same(b.graph_state(g, st), nodes.state()), "small graph %d (n = %d) vs oracle" % (g, n)
        assert lcg[g] == sim.random_source().state[0]
    for g in sorted(rng.choice(large, 3, replace=False).tolist()):
        n, src, tgt = gs[g]
        s0 = b.graph_state(g, start)
        sim, nodes = facade_run(gpu_core, n, src, tgt, 3, s0[0], s0[1])
        assert same(b.graph_state(g, st), nodes.state()), "large graph %d (n = %d) vs fast path" % (g, n)
    print("  sampled graphs identical", flush=True)
    b.close()
    # many-body error of sampled fast graphs on their starting layout
    pick = sorted(rng.choice(large, 8, replace=False).tolist())
    probe = Batch([gs[g] for g in pick], core=gpu_core)
    probe.scalar(BS_LINK_ITERATIONS, 0).scalar(BS_CENTER_STRENGTH, 0).scalar(BS_VELOCITY_DECAY, 1)
    s0 = probe.state()
    probe.tick(1)
    alpha = probe.scalar(BS_ALPHA)
    s1 = probe.state()
    worst = 0.0
    for k in range(len(pick)):
        a, c = probe.graph_state(k, s0), probe.graph_state(k, s1)
        sample = rng.choice(a.shape[1], 300, replace=False)
        fx, fy = direct(a[0], a[1], alpha, sample)
        e = math.sqrt(np.sum((c[2][sample] - fx) ** 2 + (c[3][sample] - fy) ** 2) / np.sum(fx * fx + fy * fy))
        worst = max(worst, e)
    print("  worst sampled rms vs FP64 direct sum %.2e" % worst, flush=True)
    assert worst < 1e-5
    probe.close()


def test_batch_rejects_links_across_graphs(gpu_core):
    gs = [(60, np.array([0, 1]), np.array([1, 2])), (70, np.array([0]), np.array([5]))]
    b = Batch(gs, core=gpu_core)
    src = b._u32(5, b.m)
    src[2] = 1  # graph 1's link now starts in graph 0
    with pytest.raises(Exception):
        b.sync_in()
    b.close()


def test_batch_autotune_keeps_every_bit(gpu_core, graphs):
    # The tuned launch shapes change no result: the tuning ticks and the ticks
    # after them equal an untuned batch's, bit for bit.
    a = Batch(graphs, core=gpu_core)
    b = Batch(graphs, core=gpu_core)
    info = a.autotune(1)
    print("  autotuned", info, flush=True)
    assert len([v for v in info["candidate_ms"] if v > 0]) == 14
    b.tick(15)
    assert same(a.state(), b.state())
    a.tick(3)
    b.tick(3)
    assert same(a.state(), b.state())
    a.close()
    b.close()


def test_batch_graph_replays_equal_eager_ticks(gpu_core, graphs):
    # BS_GRAPH = 1: one eager tick, then a captured tick replayed every tick.
    # Each replay is compared with an eager batch after the same tick (with
    # the host's reads and synchronisations between replays).
    from d3force_amd.batch import BS_GRAPH
    a = Batch(graphs, core=gpu_core)
    b = Batch(graphs, core=gpu_core)
    a.scalar(BS_GRAPH, 1)
    for t in range(8):
        a.tick(1)
        b.tick(1)
        assert same(a.state(), b.state()), "tick %d" % (t + 1)
    assert a.scalar(BS_ALPHA) == b.scalar(BS_ALPHA)
    a.close()
    b.close()


def test_batch_graph_redo_path_equals_eager_ticks(gpu_core, graphs):
    # With the replay's cells capped below what any tick needs
    # (BS_GRAPH_CELLS = 1), every replay aborts on the device, is undone from
    # the snapshot and redone eagerly, and the tick is captured again: the
    # path a replay takes when a layout outgrows the captured buffers.
    from d3force_amd.batch import BS_GRAPH, BS_GRAPH_CELLS, BS_GRAPH_REDOS, BS_GRAPH_REPLAYS
    a = Batch(graphs, core=gpu_core)
    b = Batch(graphs, core=gpu_core)
    a.scalar(BS_GRAPH, 1)
    a.scalar(BS_GRAPH_CELLS, 1)
    for t in range(5):
        a.tick(1)
        b.tick(1)
        assert same(a.state(), b.state()), "tick %d" % (t + 1)
    assert a.scalar(BS_ALPHA) == b.scalar(BS_ALPHA)
    assert a.scalar(BS_GRAPH_REPLAYS) == 4 and a.scalar(BS_GRAPH_REDOS) == 4
    a.close()
    b.close()
