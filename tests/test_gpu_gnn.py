"""The warm-start network's GPU kernels (BF16 matrix cores, FP32
accumulation) against the CPU restatement oracle/gnn_ref.py with BF16
rounding emulated at the same points, and against exact FP64."""

import numpy as np
import pytest

from oracle import gnn_ref as ref
from d3force_amd.warmstart import Network, csr_undirected, features, stress_pairs

pytestmark = pytest.mark.gpu


def graph(n=700, seed=2):
    rng = np.random.default_rng(seed)
    row, col = csr_undirected(n, rng.integers(0, n, 3 * n), rng.integers(0, n, 3 * n))
    return n, row, col, features(n, row, col, seed=seed)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_gnn_forward_matches_the_restatement(gpu_core):
    n, row, col, x = graph()
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
    n, row, col, x = graph(seed=4)
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
    n, row, col, x = graph(n=2000, seed=7)
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
