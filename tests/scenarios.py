"""Seeded scenario generators restated from the reference's test corpus, so
the oracle and the GPU core see identical inputs.

* ``generator``: test/parity/scenarios.js:7-16 (a 32-bit integer mixer).
* fast-mode layouts: test/threads/scenarios.js:13-56 (inputs built from
  correctly rounded arithmetic only, so the digests are portable).
* BASELINE workloads: random graph with m = 2n distinct-pair links from the
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


def ba_graph(n, m0=3, k=3, seed=1):
    """Barabasi-Albert graph: m0 fully connected seed nodes, then each new
    node v links to k distinct earlier nodes chosen with probability
    proportional to degree (uniform over the endpoints of the earlier
    links). Returns (source, target) with source = the new node. numpy
    Generator(PCG64) seeded with `seed`."""
    rng = np.random.default_rng(seed)
    s0 = [i for i in range(m0) for j in range(i + 1, m0)]
    t0 = [j for i in range(m0) for j in range(i + 1, m0)]
    e0 = len(s0)
    if n <= m0:
        keep = [(a, b) for a, b in zip(s0, t0) if b < n]
        return (np.array([a for a, _ in keep], dtype=np.int64), np.array([b for _, b in keep], dtype=np.int64))
    new = np.repeat(np.arange(m0, n, dtype=np.int64), k)
    e = np.arange(e0, e0 + len(new), dtype=np.int64)
    before = e0 + k * (new - m0)           # links that exist when node `new` arrives
    src = np.concatenate([np.array(s0, dtype=np.int64), new])
    tgt = np.empty(len(src), dtype=np.int64)
    tgt[:e0] = t0

    def resolve(r):
        # endpoint r of the link list: even = source of link r/2, odd = target
        out = np.empty(len(r), dtype=np.int64)
        idx = np.arange(len(r))
        r = r.copy()
        while len(idx):
            even = r % 2 == 0
            out[idx[even]] = src[r[even] // 2]
            idx, r = idx[~even], r[~even]
            link = (r - 1) // 2
            init = link < e0
            out[idx[init]] = tgt[link[init]]
            idx, r = idx[~init], link[~init]
            r = draws[r - e0]
        return out

    draws = (rng.random(len(new)) * (2 * before)).astype(np.int64)
    tgt[e0:] = resolve(draws)
    # Redraw repeated targets of one new node (links are distinct pairs).
    for _ in range(64):
        t = tgt[e0:].reshape(-1, k)
        dup = np.zeros(t.shape, dtype=bool)
        for a in range(k):
            for b in range(a):
                dup[:, a] |= t[:, a] == t[:, b]
        dup = dup.reshape(-1)
        if not dup.any():
            break
        which = np.nonzero(dup)[0]
        draws[which] = (rng.random(len(which)) * (2 * before[which])).astype(np.int64)
        tgt[e0 + which] = resolve(draws[which])
    return src, tgt


def sbm_batch(graphs, seed=1, n_min=50, n_max=20000, degree=6.0, p_in=0.05, p_out=0.002):
    """Stochastic block model graphs: n log-uniform in [n_min, n_max], 2-16
    equal blocks, p_in and p_out scaled by one factor per graph so that the
    expected mean degree is `degree`; distinct unordered pairs, no self-links.
    numpy Generator(PCG64) seeded with `seed`. Returns [(n, source, target)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(graphs):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        blocks = int(rng.integers(2, 17))
        block = np.minimum(np.arange(n) * blocks // n, blocks - 1)
        size = np.bincount(block, minlength=blocks).astype(np.float64)
        # expected degree of a node in block b: p_in (size_b - 1) + p_out (n - size_b)
        mean = np.sum(size * (p_in * (size - 1) + p_out * (n - size))) / n
        s = degree / mean if mean > 0 else 1.0
        pi, po = min(1.0, p_in * s), min(1.0, p_out * s)
        pairs_in = np.sum(size * (size - 1) / 2)
        pairs_out = n * (n - 1) / 2 - pairs_in
        m_in = rng.binomial(int(pairs_in), pi) if pairs_in > 0 else 0
        m_out = rng.binomial(int(pairs_out), po) if pairs_out > 0 else 0
        starts = np.concatenate([[0], np.cumsum(size)]).astype(np.int64)
        # intra-block pairs: block by weight C(size, 2), then two distinct members
        bw = size * (size - 1) / 2
        b = rng.choice(blocks, size=m_in, p=bw / bw.sum()) if m_in else np.zeros(0, dtype=np.int64)
        sz = size[b].astype(np.int64)
        u = rng.integers(0, np.maximum(sz, 1))
        v = rng.integers(0, np.maximum(sz - 1, 1))
        v = np.where(v >= u, v + 1, v)
        si, ti = starts[b] + u, starts[b] + v
        # inter-block pairs: uniform node pairs in different blocks (rejection)
        so = rng.integers(0, n, size=2 * m_out + 16)
        to = rng.integers(0, n, size=2 * m_out + 16)
        keep = block[so] != block[to]
        so, to = so[keep][:m_out], to[keep][:m_out]
        src = np.concatenate([si, so])
        tgt = np.concatenate([ti, to])
        lo, hi = np.minimum(src, tgt), np.maximum(src, tgt)
        _, first = np.unique(lo * n + hi, return_index=True)
        first.sort()
        out.append((n, src[first].astype(np.int64), tgt[first].astype(np.int64)))
    return out
