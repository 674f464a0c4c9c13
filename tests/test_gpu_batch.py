"""Batched fast mode: many graphs ticked together (d3force_amd.batch).

- A graph's result in a batch is bit-identical to a batch of that graph alone
  (nothing in a graph's arithmetic depends on the other graphs).
- A graph of at least 2000 nodes (where the facade runs fast mode) is
  bit-identical to the single-simulation GPU fast path with manyBody, link and
  center: the same FMM, link sweep, center sums and integrate (the positions
  are distinct, so no jiggle seed is drawn; seeds are the one difference
  between the two).
- The many-body force of every graph, small or large, is within FP32 rounding
  of the FMM's truncation error from an FP64 direct sum with d3's pair term.
"""

import math

import numpy as np
import pytest

import d3force_amd as d3
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
        one = Batch([graphs[g]], core=gpu_core, seed_ids=[g])
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
