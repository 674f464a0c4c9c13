"""The CPU restatement of the warm-start network (oracle/gnn_ref.py): its
analytic backward pass against central finite differences in FP64, and its
BF16 rounding against exact values."""

import numpy as np

from oracle import gnn_ref as ref
from d3force_amd.warmstart import csr_undirected


def small_case(n=40, seed=3):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, 3 * n)
    tgt = rng.integers(0, n, 3 * n)
    row, col = csr_undirected(n, src, tgt)
    x = rng.standard_normal((n, ref.F_IN))
    p = rng.standard_normal(ref.N_PARAMS) * 0.3
    w0 = p[ref.OFF_W0:ref.OFF_B0].reshape(ref.K_IN, ref.HID)
    w0[ref.F_IN:] = 0
    wo = p[ref.OFF_WO:ref.OFF_BO].reshape(ref.HID, ref.OUT_PAD)
    wo[:, ref.OUT:] = 0
    p[ref.OFF_BO + ref.OUT:] = 0
    dp = rng.standard_normal((n, 2))
    return p, x, row, col, dp


def test_backward_matches_finite_differences():
    p, x, row, col, dp = small_case()
    g = ref.backward(p, x, row, col, dp)
    rng = np.random.default_rng(0)
    live = [k for k in range(ref.N_PARAMS) if g[k] != 0]
    idx = rng.choice(live, 60, replace=False)
    h = 1e-6
    for k in idx:
        q = p.copy()
        q[k] += h
        fp = np.sum(ref.forward(q, x, row, col)[0] * dp)
        q[k] -= 2 * h
        fm = np.sum(ref.forward(q, x, row, col)[0] * dp)
        fd = (fp - fm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_padding_parameters_have_no_gradient():
    p, x, row, col, dp = small_case()
    g = ref.backward(p, x, row, col, dp)
    assert not np.any(g[ref.OFF_W0:ref.OFF_B0].reshape(ref.K_IN, ref.HID)[ref.F_IN:])
    assert not np.any(g[ref.OFF_WO:ref.OFF_BO].reshape(ref.HID, ref.OUT_PAD)[:, ref.OUT:])


def test_bf16_rounding():
    # nearest, ties to even, on values exact in BF16 and halfway between
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-40, 3.0e38], dtype=np.float32)
    r = ref.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0              # halfway: to even (1.0)
    assert r[2] == 1.0 + 2 ** -6    # halfway: to even (1 + 4 * 2^-8)
    assert r[3] == -2.5
    assert ref.bf16(np.float32(1.0 + 2 ** -7)) == 1.0 + 2 ** -7
