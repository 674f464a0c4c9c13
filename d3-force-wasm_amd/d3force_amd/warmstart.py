"""Learned warm start: a message-passing network predicts a layout, and fast
mode refines it from alpha = REFINE_ALPHA (0.5) instead of running from d3's
phyllotaxis start at alpha = 1 (the reference always starts there,
simulation.js:91-105).

The network (csrc/gnn.hip) has 8 input features per node, 4 message-passing
layers of width 64 and a 2-D position head; it runs and trains on the GPU with
BF16 matrix-core kernels. Inputs, computed on the GPU from the links alone
(gnn_set_links):
  - log(1 + degree) and 1 / sqrt(1 + degree) (degree: distinct neighbours);
  - 6 random-walk features: the smoothest non-constant modes of the lazy
    random walk (I + D^-1 A) / 2, from 40 steps of subspace iteration on the
    similar symmetric walk from seeded Gaussian vectors and Rayleigh-Ritz,
    each with mean 0, unit variance and its sign fixed by its third moment.
A prediction (gnn_predict) is the network's output centred, scaled to the size
that minimises the energy below over uniform scalings, plus a small seeded
jitter; everything from the links to the layout runs on the GPU.
It is trained on layouts the simulator itself produces at alpha ~ 0.1 (fast
mode from phyllotaxis), with the stress loss over the graph's links and random
node pairs, so the rotation and translation of the prediction are free.

Energy of a layout (the quantity the warm start is measured by; the reference
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
REFINE_ALPHA = 0.5  # where the refinement of a prediction starts (alpha decays to 0.001)


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
    "gnn_set_links": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]),
    "gnn_predict": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p]),
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

    def set_links(self, n, src, tgt, seed=0, features=False):
        """The graph from its links (rows and features on the GPU); returns
        the (n, 8) FP32 features when `features`."""
        src = np.ascontiguousarray(src, dtype=np.uint32)
        tgt = np.ascontiguousarray(tgt, dtype=np.uint32)
        f = np.empty((n, F_IN), dtype=np.float32) if features else None
        check(self.core.gnn_set_links(self.h, n, len(src), _ptr(src) if len(src) else None,
                                      _ptr(tgt) if len(tgt) else None, seed, _ptr(f) if f is not None else None))
        self.n = n
        self._keep = (src, tgt)
        return f

    def predict(self, jitter=0.02, seed=0):
        """(n, 2) FP64 layout of the graph given to set_links (module docstring)."""
        out = np.empty((self.n, 2))
        check(self.core.gnn_predict(self.h, float(jitter), seed, _ptr(out)))
        return out

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


# ---------------------------------------------------------------- pipeline

def fast_simulation(n, src, tgt, core=None, x=None, y=None):
    """The fast-mode simulation the warm start is trained on and measured by:
    manyBody + link + center with d3's defaults, FMM order 8, opening 1.5."""
    from . import forceSimulation, forceManyBody, forceLink, forceCenter, Nodes
    nodes = Nodes(n) if x is None else Nodes(n, x=x, y=y)
    sim = (forceSimulation(nodes, core=core or gpu_core())
           .force("charge", forceManyBody())
           .force("link", forceLink((src, tgt)))
           .force("center", forceCenter()).stop())
    sim.precision("fast", {"order": 8, "opening": 1.5})
    return sim, nodes


def timed_ticks(sim, core, ticks):
    """Runs `ticks` ticks through the facade (one copy in and out); returns
    their wall-clock seconds, from an idle device to the layout back on the
    host."""
    import time
    check(core.core_synchronize())
    t0 = time.perf_counter()
    sim.tick(ticks)
    check(core.core_synchronize())
    return time.perf_counter() - t0


def ticks_to_alpha(alpha0, alpha1, decay):
    """d3's alpha schedule (simulation.js:64, target 0): ticks from alpha0 to
    reach alpha1 or below."""
    return int(math.ceil(math.log(alpha1 / alpha0) / math.log(1 - decay)))


def training_graphs(count, seed=1, n_min=1000, n_max=20000):
    """Alternating Barabasi-Albert (m0 = 3, 3 links per node) and stochastic
    block model graphs with n log-uniform in [n_min, n_max]."""
    from .graphs import ba_graph, sbm_batch
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        if k % 2 == 0:
            s, t = ba_graph(n, seed=seed * 1000 + k)
            out.append((n, s, t))
        else:
            g = sbm_batch(1, seed=seed * 1000 + k, n_min=n, n_max=n)[0]
            out.append(g)
    return out


class Prepared:
    """One graph's rows, features (computed by `net` on the GPU, as a
    prediction computes them) and, for training, normalised target layout."""

    def __init__(self, n, src, tgt, net, seed=0):
        self.n, self.src, self.tgt = n, np.asarray(src), np.asarray(tgt)
        self.row, self.col = csr_undirected(n, src, tgt)
        self.seed = seed
        self.x = net.set_links(n, src, tgt, seed=seed, features=True)
        self.scale = math.sqrt(n)


def target_layout(g, core=None, alpha=0.1):
    """The simulator's own layout at `alpha` (fast mode from phyllotaxis)."""
    sim, nodes = fast_simulation(g.n, g.src, g.tgt, core=core)
    decay = 1 - math.pow(0.001, 1 / 300)
    sim.tick(ticks_to_alpha(1.0, alpha, decay))
    return np.stack([nodes.x, nodes.y], axis=1)


def train(graphs=48, epochs=30, steps=4, lr=2e-3, seed=1, out=WEIGHTS, log=print, core=None, n_max=20000,
          init=None):
    """Trains the network on simulator layouts at alpha = 0.1 of `graphs`
    training graphs (n log-uniform in [1000, n_max]) and saves its parameters
    to `out`; `init` names parameters to start from (fine-tuning) instead of
    the seeded initialisation."""
    core = core or gpu_core()
    net = Network(seed=seed, core=core)
    if init:
        net.load(init)
    gs = training_graphs(graphs, seed=seed, n_max=n_max)
    prepared = [Prepared(n, s, t, net, seed=k) for k, (n, s, t) in enumerate(gs)]
    log("warm start: features of %d graphs ready" % len(prepared))
    for k, g in enumerate(prepared):
        lay = target_layout(g, core=core)
        lay = (lay - lay.mean(axis=0)) / g.scale
        g.pairs = stress_pairs(g.row, g.col, lay, seed=k)
        if (k + 1) % 100 == 0:
            log("warm start: %d training layouts ready" % (k + 1))
    log("warm start: %d training layouts ready" % len(prepared))
    rng = np.random.default_rng(seed)
    total = epochs * len(prepared) * steps
    done = 0
    for e in range(epochs):
        losses = []
        for k in rng.permutation(len(prepared)):
            g = prepared[k]
            net.set_graph(g.row, g.col, g.x)
            net.set_pairs(*g.pairs)
            for _ in range(steps):
                rate = lr * 0.5 * (1 + math.cos(math.pi * done / total))
                losses.append(net.train_step(rate))
                done += 1
        log("warm start: epoch %d stress %.4f" % (e + 1, float(np.mean(losses))))
    if out:
        net.save(out)
    return net


def predict(net, n, src, tgt, jitter=0.02, seed=0):
    """Predicted layout (n x 2) in the simulation's units, all on the GPU:
    rows and features from the links, the network's shape centred and scaled
    to the energy-minimising size (the network's global scale is its least
    reliable output, and a layout much larger or smaller than the
    simulation's costs refinement ticks), plus a seeded uniform jitter of
    +-`jitter` times the mean predicted link length. Nodes with identical
    features and neighbourhoods (leaves of one hub, isolated nodes) get
    identical predictions; the jitter spreads them on the layout's own
    scale."""
    net.set_links(n, src, tgt, seed=seed)
    return net.predict(jitter=jitter, seed=seed)


def graph_stress(x, y, row, col, sources=16, seed=5):
    """Sampled graph-distance stress of a layout: over the pairs (source,
    node) for `sources` random sources and every node they reach,
    sum ((|p_i - p_j| - c D_ij) / (c D_ij))^2 / pairs with D the hop
    distance and c the scale that minimises it."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import shortest_path
    n = len(row) - 1
    rows = np.repeat(np.arange(n), np.diff(row.astype(np.int64)))
    a = sp.csr_matrix((np.ones(len(col)), (rows, col.astype(np.int64))), shape=(n, n))
    src = np.random.default_rng(seed).choice(n, min(n, sources), replace=False)
    dist = shortest_path(a, unweighted=True, indices=src)
    pairs = []
    for k, s_ in enumerate(src):
        d = dist[k]
        ok = np.isfinite(d) & (d > 0)
        r = np.hypot(x[ok] - x[s_], y[ok] - y[s_])
        pairs.append((r, d[ok]))
    # c minimising sum (r / (c D) - 1)^2: c = sum (r / D)^2 / sum (r / D)
    q = np.concatenate([r / d for r, d in pairs])
    c = np.sum(q * q) / np.sum(q)
    return float(np.mean((q / c - 1) ** 2))


def compare(net, n, src, tgt, core=None, seed=0, short=(50, 100), stress_all=True):
    """Plain fast mode from phyllotaxis (300 ticks, d3's schedule) against
    the warm start (the prediction refined from alpha = REFINE_ALPHA, alpha
    decaying to 0.001 over K ticks) and against plain fast mode squeezed into the same K
    ticks (alpha 1 -> 0.001 over K ticks: the method without the network),
    for each K of `short`. All runs reuse one simulation (its buffers are
    allocated by a first, untimed run). Seconds are wall-clock on both sides:
    a plain run from its start layout on the host to its layout back on the
    host; a warm run from the links on the host (rows, features, network,
    scaling and jitter on the GPU, the layout to the host) through its
    refinement ticks to the layout back on the host.
    Energies use one sample of node pairs; `gap` is (E - E_plain) /
    (E_start - E_plain): 0 is as good as 300 plain ticks, 1 no better than
    the start. With stress_all False the squeezed plain runs skip the
    (host-side) stress measurement."""
    import time
    core = core or gpu_core()
    src, tgt = np.asarray(src), np.asarray(tgt)
    row, col = csr_undirected(n, src, tgt)
    k_link = _link_strength(n, src, tgt)
    sim, nodes = fast_simulation(n, src, tgt, core=core)
    x0, y0 = nodes.x.copy(), nodes.y.copy()
    plain_decay = sim.alpha_decay()

    def run(x, y, alpha, decay, ticks):
        nodes.x[:] = x
        nodes.y[:] = y
        nodes.vx[:] = 0
        nodes.vy[:] = 0
        sim.alpha(alpha).alpha_decay(decay)
        return timed_ticks(sim, core, ticks)

    run(x0, y0, 1.0, plain_decay, 5)  # allocations, untimed
    e_start, pairs = energy(x0, y0, src, tgt, k_link)
    t_plain = run(x0, y0, 1.0, plain_decay, 300)
    e_plain, _ = energy(nodes.x, nodes.y, src, tgt, k_link, pairs=pairs)
    span = e_start - e_plain
    out = {"nodes": n, "links": len(src), "start_energy": e_start, "plain_energy": e_plain, "plain_ticks": 300,
           "plain_seconds": t_plain, "plain_stress": graph_stress(nodes.x, nodes.y, row, col), "runs": []}
    predict(net, n, src, tgt, seed=seed)  # first call allocates
    check(core.core_synchronize())
    t0 = time.perf_counter()
    p = predict(net, n, src, tgt, seed=seed)
    t_predict = time.perf_counter() - t0
    e_pred, _ = energy(p[:, 0], p[:, 1], src, tgt, k_link, pairs=pairs)
    out["predict_seconds"] = t_predict
    out["predicted_energy_gap"] = (e_pred - e_plain) / span
    a0 = REFINE_ALPHA
    run(p[:, 0], p[:, 1], a0, 1 - math.pow(0.001 / a0, 1 / max(short)), 3)  # buffers for this layout, untimed
    for k in short:
        t_warm = t_predict + run(p[:, 0], p[:, 1], a0, 1 - math.pow(0.001 / a0, 1 / k), k)
        e_warm, _ = energy(nodes.x, nodes.y, src, tgt, k_link, pairs=pairs)
        s_warm = graph_stress(nodes.x, nodes.y, row, col)
        t_short = run(x0, y0, 1.0, 1 - math.pow(0.001, 1 / k), k)
        e_short, _ = energy(nodes.x, nodes.y, src, tgt, k_link, pairs=pairs)
        s_short = graph_stress(nodes.x, nodes.y, row, col) if stress_all else None
        out["runs"].append({
            "ticks": k,
            "warm": {"energy": e_warm, "gap": (e_warm - e_plain) / span,
                     "relative_to_plain": (e_warm - e_plain) / abs(e_plain), "seconds": t_warm, "stress": s_warm},
            "plain_same_ticks": {"energy": e_short, "gap": (e_short - e_plain) / span,
                                 "relative_to_plain": (e_short - e_plain) / abs(e_plain), "seconds": t_short,
                                 "stress": s_short}})
    return out


def heldout_graph(k, nodes=100000, seed=20000):
    """Held-out graph k of the evaluation: stochastic block model for even k,
    Barabasi-Albert (m0 = 3, 3 links per node) for odd k, seeds disjoint from
    the training graphs'."""
    from .graphs import ba_graph, sbm_batch
    if k % 2 == 0:
        return "sbm", sbm_batch(1, seed=seed + k, n_min=nodes, n_max=nodes)[0]
    return "ba", (nodes,) + ba_graph(nodes, seed=seed + k)


def evaluate(net, count, nodes=100000, short=(25, 50, 75), core=None, log=print):
    """The warm start against 300 plain fast-mode ticks on `count` held-out
    graphs (heldout_graph): per graph, the first K of `short` whose warm run
    ends within 1% of the plain energy and its GPU-time ratio, the energy
    gaps at equal ticks and the stresses; then a summary per graph family."""
    rows = []
    for k in range(count):
        family, (n, src, tgt) = heldout_graph(k, nodes)
        r = compare(net, n, src, tgt, core=core, seed=k, short=short, stress_all=False)
        hit = next((run for run in r["runs"] if abs(run["warm"]["relative_to_plain"]) <= 0.01), None)
        row = {"graph": k, "family": family, "nodes": n, "links": len(src),
               "plain_seconds": r["plain_seconds"], "plain_stress": r["plain_stress"],
               "predicted_energy_gap": r["predicted_energy_gap"],
               "hit_ticks": hit["ticks"] if hit else None,
               "speedup": r["plain_seconds"] / hit["warm"]["seconds"] if hit else None,
               "hit_stress": hit["warm"]["stress"] if hit else None,
               "gaps": {run["ticks"]: (run["warm"]["gap"], run["plain_same_ticks"]["gap"]) for run in r["runs"]}}
        rows.append(row)
        log("held-out %d (%s): within 1%% at %s ticks, speedup %s, predicted gap %.3f, gaps %s"
            % (k, family, row["hit_ticks"], None if hit is None else round(row["speedup"], 2),
               row["predicted_energy_gap"],
               {t: (round(a, 4), round(b, 4)) for t, (a, b) in row["gaps"].items()}))
    summary = {}
    for family in sorted(set(r["family"] for r in rows)):
        fr = [r for r in rows if r["family"] == family]
        sp = sorted(r["speedup"] for r in fr if r["speedup"] is not None)
        mid = short[len(short) // 2]
        summary[family] = {
            "graphs": len(fr), "within_1pct": len(sp),
            "speedup_min": sp[0] if sp else None, "speedup_median": sp[len(sp) // 2] if sp else None,
            "speedup_max": sp[-1] if sp else None,
            "at_%d_ticks_warm_gap_median" % mid: float(np.median([r["gaps"][mid][0] for r in fr])),
            "at_%d_ticks_plain_gap_median" % mid: float(np.median([r["gaps"][mid][1] for r in fr])),
            "plain_stress_median": float(np.median([r["plain_stress"] for r in fr])),
            "hit_stress_median": float(np.median([r["hit_stress"] for r in fr if r["hit_stress"] is not None]))
            if sp else None}
        log("summary %s: %s" % (family, summary[family]))
    return {"graphs": rows, "summary": summary}


def _link_strength(n, src, tgt):
    count = np.bincount(src, minlength=n) + np.bincount(tgt, minlength=n)
    return 1.0 / np.minimum(count[src], count[tgt])


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="train the warm-start network on the GPU")
    ap.add_argument("--graphs", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--steps", type=int, default=4, help="Adam steps per graph and epoch")
    ap.add_argument("--n-max", type=int, default=20000, help="largest training graph (nodes)")
    ap.add_argument("--lr", type=float, default=2e-3, help="peak Adam learning rate (cosine schedule)")
    ap.add_argument("--init", default=None, help="parameters to fine-tune from")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--evaluate", type=int, default=0,
                    help="instead of training: evaluate the committed network on this many held-out graphs")
    ap.add_argument("--nodes", type=int, default=100000, help="nodes per held-out graph (with --evaluate)")
    ap.add_argument("--weights", default=WEIGHTS, help="network parameters to evaluate (with --evaluate)")
    a = ap.parse_args()
    if a.evaluate:
        import json
        net = Network()
        net.load(a.weights)
        res = evaluate(net, a.evaluate, nodes=a.nodes, log=lambda *m: print(*m, flush=True))
        net.close()
        print(json.dumps(res["summary"]))
    else:
        train(graphs=a.graphs, epochs=a.epochs, steps=a.steps, lr=a.lr, out=a.out, n_max=a.n_max,
              init=a.init, log=lambda *m: print(*m, flush=True))
