"""Seeded scenario generators restated from the reference's test corpus, so
the oracle and the GPU core see identical inputs.

* ``generator``: test/parity/scenarios.js:7-16 (a 32-bit integer mixer).
* fast-mode layouts: test/threads/scenarios.js:13-56 (inputs built from
  correctly rounded arithmetic only, so the digests are portable).
* benchmark workloads: random graph with m = 2n distinct-pair links from the
  d3 LCG, phyllotaxis start, optional pinned nodes.
"""

import hashlib
import math

import numpy as np

M32 = 0xFFFFFFFF


def generator(seed):
    """test/parity/scenarios.js:7-16."""
    state = [seed & M32]

    def imul(a, b):
        return ((a & M32) * (b & M32)) & M32

    def r():
        state[0] = (state[0] + 0x6D2B79F5) & M32
        t = state[0]
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & M32
        return ((t ^ (t >> 14)) & M32) / 4294967296

    return r


def disk(r, cx, cy, radius):
    """test/threads/scenarios.js:13-17."""
    while True:
        x = 2 * r() - 1
        y = 2 * r() - 1
        if not (x * x + y * y > 1):
            return cx + radius * x, cy + radius * y


def spiral(n):
    """test/threads/scenarios.js:21-29: phyllotaxis by a fixed rotation."""
    c, s = -0.7373688780783197, 0.6754902942615238
    xs, ys = np.empty(n), np.empty(n)
    u, v = 1.0, 0.0
    for i in range(n):
        radius = 10 * math.sqrt(0.5 + i)
        xs[i] = radius * u
        ys[i] = radius * v
        u, v = u * c - v * s, u * s + v * c
    return xs, ys


def _pairs(pts):
    xs = np.array([p[0] for p in pts], dtype=np.float64)
    ys = np.array([p[1] for p in pts], dtype=np.float64)
    return xs, ys


FAST_LAYOUTS = {
    "uniform 4000": lambda r: _pairs([(1000 * (r() - 0.5), 1000 * (r() - 0.5)) for _ in range(4000)]),
    "three clusters 6000": lambda r: _pairs([disk(r, [0, 400, -300][i % 3], [0, 100, 350][i % 3], 120)
                                             for i in range(6000)]),
    "spiral 8000": lambda r: spiral(8000),
    "grid 2500 with exact zero deltas": lambda r: _pairs([(12 * (i % 50), 12 * (i // 50)) for i in range(2500)]),
    "dense sub-unit cluster 3000": lambda r: _pairs([(20 * r(), 20 * r()) for _ in range(3000)]),
}

CHECKPOINTS = [1, 10, 100, 300]


def fast_scenario(name, seed=1):
    """(xs, ys, source, target) of test/threads/scenarios.js:43-56."""
    r = generator(seed)
    xs, ys = FAST_LAYOUTS[name](r)
    n = len(xs)
    src = np.arange(n, dtype=np.int64)
    tgt = np.array([math.floor(r() * n) for _ in range(n)], dtype=np.int64)
    return xs, ys, src, tgt


def digest(nodes):
    """test/threads/digest.js:1-13: SHA-256 of x, y, vx, vy per node as f64."""
    a = np.stack([nodes.x, nodes.y, nodes.vx, nodes.vy], axis=1).astype("<f8")
    return hashlib.sha256(a.tobytes()).hexdigest()


def lcg_stream(seed=1):
    """d3's LCG (src/lcg.js) as a generator for data, seeded."""
    s = seed

    def r():
        nonlocal s
        s = (1664525 * s + 1013904223) % 4294967296
        return s / 4294967296

    return r


def random_graph(n, m, seed=1):
    """m links over distinct unordered node pairs, endpoints floor(r n) from a
    seeded d3 LCG; self-pairs and repeats are rejected and redrawn."""
    r = lcg_stream(seed)
    # Draw in batches with numpy for speed while keeping the LCG stream.
    src = np.empty(m, dtype=np.int64)
    tgt = np.empty(m, dtype=np.int64)
    seen = set()
    k = 0
    a, c, mod = 1664525, 1013904223, 4294967296
    s = seed
    while k < m:
        s = (a * s + c) % mod
        i = (s * n) // mod
        s = (a * s + c) % mod
        j = (s * n) // mod
        if i == j:
            continue
        key = (i, j) if i < j else (j, i)
        if key in seen:
            continue
        seen.add(key)
        src[k] = i
        tgt[k] = j
        k += 1
    del r
    return src, tgt


def pinned(n, fraction, seed=7):
    """Indices and (fx, fy) in [-500, 500] for ``fraction`` of the nodes,
    chosen by a seeded d3 LCG."""
    r = lcg_stream(seed)
    want = int(round(n * fraction))
    chosen = []
    seen = set()
    while len(chosen) < want:
        i = int(r() * n)
        if i in seen:
            continue
        seen.add(i)
        chosen.append(i)
    idx = np.array(sorted(chosen), dtype=np.int64)
    fx = np.array([1000 * (r() - 0.5) for _ in idx])
    fy = np.array([1000 * (r() - 0.5) for _ in idx])
    return idx, fx, fy


from d3force_amd.graphs import ba_graph, sbm_batch  # noqa: E402,F401  (shared with the benchmark)
