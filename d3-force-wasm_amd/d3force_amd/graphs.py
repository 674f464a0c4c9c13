"""Synthetic graph families of the benchmark workloads: Barabasi-Albert graphs
and stochastic block models, generated with numpy from a seed."""

import math

import numpy as np


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
