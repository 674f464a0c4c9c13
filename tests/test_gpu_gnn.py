"""The warm-start network's GPU kernels (BF16 matrix cores, FP32
accumulation) against the CPU restatement oracle/gnn_ref.py with BF16
rounding emulated at the same points, and against exact FP64; and its
inputs, computed on the GPU from the links (gnn_set_links, gnn_predict),
against what the mathematics fixes: the rows equal the host's symmetric
rows, the degree features their formula, the mode features are normalised
and span the known smooth modes of a cycle and the block split of a
two-block graph, and the predicted scale is a stationary point of d3's
energy over uniform scalings."""

import math

import numpy as np
import pytest

from oracle import gnn_ref as ref
from d3force_amd.warmstart import Network, csr_undirected, stress_pairs, _link_strength
from d3force_amd.graphs import sbm_batch

pytestmark = pytest.mark.gpu


def links(n=700, seed=2):
    rng = np.random.default_rng(seed)
    return rng.integers(0, n, 3 * n), rng.integers(0, n, 3 * n)


def graph(core, n=700, seed=2):
    """Rows (host) and GPU features of a random graph with repeated links and
    self-links."""
    src, tgt = links(n, seed)
    row, col = csr_undirected(n, src, tgt)
    net = Network(seed=1, core=core)
    x = net.set_links(n, src, tgt, seed=seed, features=True)
    net.close()
    return n, row, col, x


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_gnn_forward_matches_the_restatement(gpu_core):
    n, row, col, x = graph(gpu_core)
    net = Network(seed=5, core=gpu_core)
    net.set_graph(row, col, x)
    got = net.forward()
    p = net.params()
    want_bf, _ = ref.forward(p, x, row, col, bf=True)
    want, _ = ref.forward(p, x, row, col, bf=False)
    print("  forward: vs BF16 restatement %.2e, vs FP64 %.2e" % (rel(got, want_bf), rel(got, want)), flush=True)
    assert rel(got, want_bf) < 1e-2
    assert rel(got, want) < 5e-2
    net.close()


def test_gnn_gradient_matches_the_restatement(gpu_core):
    n, row, col, x = graph(gpu_core, seed=4)
    net = Network(seed=6, core=gpu_core)
    net.set_graph(row, col, x)
    dp = np.random.default_rng(0).standard_normal((n, 2)).astype(np.float32)
    got = net.gradient(dp)
    p = net.params()
    want_bf = ref.backward(p, x, row, col, dp, bf=True)
    want = ref.backward(p, x, row, col, dp, bf=False)
    cos = float(np.dot(got, want) / (np.linalg.norm(got) * np.linalg.norm(want)))
    print("  gradient: vs BF16 restatement %.2e, vs FP64 %.2e (cosine %.4f)" % (rel(got, want_bf), rel(got, want), cos),
          flush=True)
    # BF16 activations and gradients through five layers: the exact FP64
    # gradient is matched in direction, the BF16 restatement closely.
    assert rel(got, want_bf) < 3e-2
    assert cos > 0.99
    net.close()


def test_gnn_training_lowers_the_stress(gpu_core):
    n, row, col, x = graph(gpu_core, n=2000, seed=7)
    rng = np.random.default_rng(3)
    target = rng.standard_normal((n, 2)) * 100 + np.stack([x[:, 2], x[:, 3]], axis=1) * 300
    pi, pj, d = stress_pairs(row, col, target, seed=1)
    net = Network(seed=8, core=gpu_core)
    net.set_graph(row, col, x)
    net.set_pairs(pi, pj, d)
    losses = [net.train_step(3e-3) for _ in range(60)]
    print("  stress %.3f -> %.3f" % (losses[0], losses[-1]), flush=True)
    assert losses[-1] < 0.5 * losses[0]
    net.close()


def test_set_links_rows_equal_the_host_rows(gpu_core):
    # The forward pass after gnn_set_links (rows built on the GPU) equals, bit
    # for bit, the forward pass after gnn_set_graph with the host's symmetric
    # rows and the same features: the neighbour means read the same rows.
    print("gnn inputs: GPU rows vs host rows", flush=True)
    n = 3000
    src, tgt = links(n, 11)
    src[:50] = tgt[:50]  # self-links
    src[50:100], tgt[50:100] = src[100:150], tgt[100:150]  # repeats
    src[150:160], tgt[150:160] = tgt[100:110], src[100:110]  # repeats reversed
    net = Network(seed=3, core=gpu_core)
    x = net.set_links(n, src, tgt, seed=5, features=True)
    a = net.forward()
    row, col = csr_undirected(n, src, tgt)
    net.set_graph(row, col, x)
    b = net.forward()
    assert np.array_equal(a, b)
    x2 = net.set_links(n, src, tgt, seed=5, features=True)
    assert np.array_equal(x, x2)  # deterministic
    net.close()


def test_feature_columns(gpu_core):
    print("gnn inputs: degree features and mode normalisation", flush=True)
    n = 5000
    src, tgt = links(n, 12)
    src[:300] = 7  # a hub
    keep = np.ones(n, bool)
    keep[4000:4100] = False  # isolated nodes
    ok = keep[src] & keep[tgt]
    src, tgt = src[ok], tgt[ok]
    row, col = csr_undirected(n, src, tgt)
    deg = np.diff(row.astype(np.int64)).astype(np.float64)
    net = Network(seed=3, core=gpu_core)
    x = net.set_links(n, src, tgt, seed=9, features=True)
    net.close()
    assert np.allclose(x[:, 0], np.log1p(deg).astype(np.float32), rtol=2e-7, atol=0)
    assert np.allclose(x[:, 1], (1 / np.sqrt(1 + deg)).astype(np.float32), rtol=2e-7, atol=0)
    f = x[:, 2:].astype(np.float64)
    assert np.all(np.abs(f.mean(axis=0)) < 1e-5)
    assert np.allclose(f.std(axis=0), 1, atol=1e-5)
    assert np.all(np.mean(f ** 3, axis=0) >= 0)
    assert np.all(f[4000:4100] == f[4000])  # isolated nodes: no mode reaches them


def test_cycle_modes_span_the_first_frequencies(gpu_core):
    # The lazy walk on a cycle of 24 nodes has the modes cos and sin of
    # 2 pi k i / 24 with eigenvalues (1 + cos(2 pi k / 24)) / 2: the first two
    # features span frequency 1, the next two frequency 2 (40 steps leave
    # errors of about (0.75 / 0.933)^40 = 2e-4 there).
    print("gnn inputs: cycle modes", flush=True)
    n = 24
    i = np.arange(n)
    net = Network(seed=3, core=gpu_core)
    x = net.set_links(n, i, (i + 1) % n, seed=2, features=True)
    net.close()
    for freq, cols in ((1, (2, 3)), (2, (4, 5))):
        basis = np.stack([np.cos(2 * np.pi * freq * i / n), np.sin(2 * np.pi * freq * i / n)], axis=1)
        q, _ = np.linalg.qr(basis)
        for k in cols:
            f = x[:, k].astype(np.float64)
            assert np.linalg.norm(f - q @ (q.T @ f)) < 2e-3 * np.linalg.norm(f), (freq, k)


def test_two_block_graph_first_mode_splits_the_blocks(gpu_core):
    print("gnn inputs: two-block split", flush=True)
    rng = np.random.default_rng(4)
    n = 4000
    half = n // 2
    a = rng.integers(0, half, 12000)
    b = rng.integers(0, half, 12000)
    s_ = np.concatenate([a, a + half, rng.integers(0, half, 60)])
    t_ = np.concatenate([b, b + half, rng.integers(half, n, 60)])
    net = Network(seed=3, core=gpu_core)
    x = net.set_links(n, s_, t_, seed=6, features=True)
    net.close()
    side = x[:, 2] > 0
    agree = max(np.mean(side[:half]) + np.mean(~side[half:]), np.mean(~side[:half]) + np.mean(side[half:])) / 2
    assert agree > 0.98


def test_predicted_scale_is_stationary_for_the_energy(gpu_core):
    # gnn_predict without jitter: centred, and its size s minimises
    # E(s) = sum k (s r - 30)^2 / 2 - 30 n (n - 1) / 2 log(s) (d3's link
    # and pair potential under a uniform scaling): dE/ds = 0 at the
    # predicted layout, checked by central differences.
    print("gnn inputs: predicted scale", flush=True)
    n, src, tgt = sbm_batch(1, seed=31, n_min=6000, n_max=6000)[0]
    net = Network(seed=3, core=gpu_core)
    net.set_links(n, src, tgt, seed=1)
    p = net.predict(jitter=0.0, seed=1)
    net.close()
    assert np.all(np.abs(p.mean(axis=0)) < 1e-6 * np.abs(p).max())
    k = _link_strength(n, src, tgt)
    r = np.hypot(*(p[src] - p[tgt]).T)
    c = 30.0 * n * (n - 1) / 2

    def energy(s):
        return np.sum(k * (s * r - 30) ** 2) / 2 - c * math.log(s)
    h = 1e-5
    slope = (energy(1 + h) - energy(1 - h)) / (2 * h)
    curvature = (energy(1 + h) - 2 * energy(1) + energy(1 - h)) / (h * h)
    assert abs(slope) < 1e-6 * curvature
