"""Learned warm start: a message-passing network predicts a layout, and fast
mode refines it from alpha = 0.1 instead of running from d3's phyllotaxis
start at alpha = 1 (the reference always starts there, simulation.js:91-105).

The network (csrc/gnn.hip) has 8 input features per node, 4 message-passing
layers of width 64 and a 2-D position head; it runs and trains on the GPU with
BF16 matrix-core kernels. Inputs:
  - log(1 + degree) and 1 / sqrt(1 + degree);
  - 6 random-walk features: Ritz vectors of the lazy random walk
    (I + D^-1 A) / 2 after subspace iteration from seeded Gaussian vectors
    (the graph's smoothest non-constant modes), each with unit variance and
    its sign fixed by its third moment.
It is trained on layouts the simulator itself produces at alpha ~ 0.1 (fast
mode from phyllotaxis), with the stress loss over the graph's links and random
node pairs, so the rotation and translation of the prediction are free.

Energy of a layout (the quantity the warm start is judged by; the reference
defines none): the potential of d3's forces with their default parameters,
  E = sum over links 1/2 k (r - d)^2 + sum over node pairs |s| log(1 / r'),
with k the link strength, d = 30, s = -30 and r' = max(r, 1) (distanceMin
1); the pair sum is estimated from a fixed sample of node pairs.
"""

import ctypes
import math
import os

import numpy as np

from .core import check, gpu_core

F_IN = 8


def csr_undirected(n, src, tgt):
    """Symmetric adjacency rows without self-links or repeated neighbours."""
    src = np.asarray(src, dtype=np.int64)
    tgt = np.asarray(tgt, dtype=np.int64)
    keep = src != tgt
    a = np.concatenate([src[keep], tgt[keep]])
    b = np.concatenate([tgt[keep], src[keep]])
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    row = np.zeros(n + 1, dtype=np.uint32)
    np.add.at(row, a + 1, 1)
    row = np.cumsum(row).astype(np.uint32)
    return row, b.astype(np.uint32)


def features(n, row, col, seed=0, modes=6, iterations=40):
    """(n, 8) float32 node features (see the module docstring)."""
    import scipy.sparse as sp
    deg = np.diff(row.astype(np.int64)).astype(np.float64)
    f = np.zeros((n, F_IN), dtype=np.float64)
    f[:, 0] = np.log1p(deg)
    f[:, 1] = 1.0 / np.sqrt(1.0 + deg)
    if n > modes + 1:
        data = np.ones(len(col))
        rows = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
        a = sp.csr_matrix((data, (rows, col.astype(np.int64))), shape=(n, n))
        dinv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
        walk = sp.diags(dinv) @ a
        rng = np.random.default_rng(seed)
        x = rng.standard_normal((n, modes + 1))
        for it in range(iterations):
            x = 0.5 * (x + walk @ x)
            if it % 5 == 4 or it == iterations - 1:
                x -= x.mean(axis=0)
                x, _ = np.linalg.qr(x)
        # Rayleigh-Ritz on the lazy walk, smoothest modes first
        mx = 0.5 * (x + walk @ x)
        h = x.T @ mx
        w, v = np.linalg.eig(h)
        order = np.argsort(-w.real)
        ritz = (x @ v[:, order].real)[:, :modes]
        for k in range(modes):
            c = ritz[:, k] - ritz[:, k].mean()
            sd = c.std()
            c = c / sd if sd > 0 else c
            if np.mean(c ** 3) < 0:
                c = -c
            f[:, 2 + k] = c
    return f.astype(np.float32)


def stress_pairs(row, col, layout, random_pairs=4, seed=0):
    """(i, j, d) over every link (both directions once) and `random_pairs`
    random partners per node, d the distance in `layout` (pairs at distance
    0 dropped)."""
    n = len(row) - 1
    rows = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    keep = rows < col
    pi = [rows[keep], ]
    pj = [col[keep].astype(np.int64), ]
    rng = np.random.default_rng(seed)
    for _ in range(random_pairs):
        pi.append(np.arange(n))
        pj.append(rng.integers(0, n, n))
    pi = np.concatenate(pi)
    pj = np.concatenate(pj)
    d = np.hypot(layout[pi, 0] - layout[pj, 0], layout[pi, 1] - layout[pj, 1])
    ok = (pi != pj) & (d > 0) & np.isfinite(d)
    return pi[ok].astype(np.uint32), pj[ok].astype(np.uint32), d[ok].astype(np.float32)


def energy(x, y, src, tgt, link_strength, pairs=None, samples=1_000_000, seed=11, charge=-30.0, distance=30.0):
    """The layout energy of the module docstring; `pairs` (i, j) fixes the
    sample of node pairs (drawn from `seed` when None)."""
    n = len(x)
    dx = x[tgt] - x[src]
    dy = y[tgt] - y[src]
    r = np.hypot(dx, dy)
    e_link = 0.5 * np.sum(link_strength * (r - distance) ** 2)
    if pairs is None:
        rng = np.random.default_rng(seed)
        i = rng.integers(0, n, samples)
        j = rng.integers(0, n, samples)
        ok = i != j
        pairs = (i[ok], j[ok])
    i, j = pairs
    rr = np.maximum(np.hypot(x[i] - x[j], y[i] - y[j]), 1.0)
    e_pair = abs(charge) * np.mean(-np.log(rr)) * n * (n - 1) / 2
    return e_link + e_pair, pairs


_SIGNATURES = {
    "gnn_create": (ctypes.c_uint32, [ctypes.c_uint64]),
    "gnn_destroy": (ctypes.c_int32, [ctypes.c_uint32]),
    "gnn_param_count": (ctypes.c_uint32, []),
    "gnn_params": (ctypes.c_void_p, [ctypes.c_uint32]),
    "gnn_sync_params": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "gnn_set_graph": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "gnn_forward": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_void_p]),
    "gnn_set_pairs": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "gnn_gradient": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    "gnn_train_step": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_float, ctypes.c_void_p]),
}


def _bind(core):
    if not hasattr(core, "gnn_create"):
        for name, (res, args) in _SIGNATURES.items():
            setattr(core, name, core.bind(name, res, args))
    return core


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Network:
    """The warm-start network on the GPU (csrc/gnn.hip)."""

    def __init__(self, seed=1, core=None):
        self.core = _bind(core or gpu_core())
        self.h = self.core.gnn_create(seed)
        if not self.h:
            raise RuntimeError("gnn_create failed (no GPU?)")
        self.size = self.core.gnn_param_count()
        self.n = 0

    def params(self):
        """The FP32 master parameters (a copy)."""
        check(self.core.gnn_sync_params(self.h, 1))
        p = ctypes.cast(self.core.gnn_params(self.h), ctypes.POINTER(ctypes.c_float))
        return np.ctypeslib.as_array(p, shape=(self.size,)).copy()

    def set_params(self, values):
        p = ctypes.cast(self.core.gnn_params(self.h), ctypes.POINTER(ctypes.c_float))
        np.ctypeslib.as_array(p, shape=(self.size,))[:] = np.asarray(values, dtype=np.float32)
        check(self.core.gnn_sync_params(self.h, 0))

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def set_graph(self, row, col, feats):
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        feats = np.ascontiguousarray(feats, dtype=np.float32)
        self.n = len(row) - 1
        check(self.core.gnn_set_graph(self.h, self.n, _ptr(row), _ptr(col) if len(col) else None, _ptr(feats)))
        self._keep = (row, col, feats)

    def forward(self):
        out = np.empty((self.n, 2), dtype=np.float32)
        check(self.core.gnn_forward(self.h, _ptr(out)))
        return out

    def gradient(self, dp):
        dp = np.ascontiguousarray(dp, dtype=np.float32)
        g = np.empty(self.size, dtype=np.float32)
        check(self.core.gnn_gradient(self.h, _ptr(dp), _ptr(g)))
        return g

    def set_pairs(self, pi, pj, d):
        pi = np.ascontiguousarray(pi, dtype=np.uint32)
        pj = np.ascontiguousarray(pj, dtype=np.uint32)
        d = np.ascontiguousarray(d, dtype=np.float32)
        check(self.core.gnn_set_pairs(self.h, len(pi), _ptr(pi), _ptr(pj), _ptr(d)))
        self._pairs = (pi, pj, d)

    def train_step(self, lr):
        loss = np.zeros(1, dtype=np.float32)
        check(self.core.gnn_train_step(self.h, float(lr), _ptr(loss)))
        return float(loss[0])

    def close(self):
        if self.h:
            self.core.gnn_destroy(self.h)
            self.h = 0


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "warmstart_weights.npy")
