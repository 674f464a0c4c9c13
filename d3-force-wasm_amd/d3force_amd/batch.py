"""Batched fast mode: many independent d3 simulations ticked together.

The reference runs one simulation per core instance; here every graph of a
batch shares one node arena and one link arena in HBM and one launch sequence
per tick (include/d3fgpu.h, batch_*). Per graph, a tick is the facade's fast
tick with forceManyBody (FMM), forceLink and forceCenter in that order, then
integrate, with d3's defaults (manyBody.js:5-12, link.js:30-32 and 56-58,
center.js:4-7, simulation.js:18-31).
"""

import ctypes
import math

import numpy as np

from .core import check, gpu_core

B_X, B_Y, B_VX, B_VY, B_CHARGE, B_SOURCE, B_TARGET, B_LINK_STRENGTH, B_LINK_DISTANCE, B_LINK_BIAS, B_SEED = range(11)
(BS_ALPHA, BS_ALPHA_DECAY, BS_ALPHA_TARGET, BS_VELOCITY_DECAY, BS_DISTANCE_MIN2, BS_DISTANCE_MAX2, BS_ORDER,
 BS_OPENING, BS_CENTER_X, BS_CENTER_Y, BS_CENTER_STRENGTH, BS_LINK_ITERATIONS, BS_TICK, BS_GRAPH,
 BS_GRAPH_CELLS, BS_GRAPH_REPLAYS, BS_GRAPH_REDOS, BS_THETA2) = range(18)
MIN_N = 2000  # fmm.zig:43: graphs below this many nodes run the exact kernels

_SIGNATURES = {
    "batch_create": (ctypes.c_uint32, []),
    "batch_destroy": (ctypes.c_int32, [ctypes.c_uint32]),
    "batch_set_graphs": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
    "batch_ptr": (ctypes.c_void_p, [ctypes.c_uint32, ctypes.c_uint32]),
    "batch_sync": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "batch_set_scalar": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double]),
    "batch_get_scalar": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32]),
    "batch_tick": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "batch_set_timing": (None, [ctypes.c_int]),
    "batch_autotune": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]),
}


def _bind(core):
    if not hasattr(core, "batch_create"):
        for name, (res, args) in _SIGNATURES.items():
            setattr(core, name, core.bind(name, res, args))
        from .core import Stats
        core.batch_stats = core.bind("batch_stats", ctypes.c_int32, [ctypes.c_uint32, ctypes.POINTER(Stats)])
    return core


def phyllotaxis_local(n):
    """d3's initial layout (simulation.js:96-99) for nodes 0..n-1 of one graph."""
    i = np.arange(n, dtype=np.float64)
    r = 10.0 * np.sqrt(0.5 + i)
    a = i * (math.pi * (3 - math.sqrt(5)))
    return r * np.cos(a), r * np.sin(a)


def link_defaults(n, src, tgt):
    """link.js:56-58 and 30-32: bias = count(s) / (count(s) + count(t)),
    strength = 1 / min(count(s), count(t)), with count the node degrees."""
    count = (np.bincount(src, minlength=n) + np.bincount(tgt, minlength=n)).astype(np.float64)
    bias = count[src] / (count[src] + count[tgt])
    strength = 1.0 / np.minimum(count[src], count[tgt])
    return strength, bias


class Batch:
    """graphs: a list of (n, source, target) with graph-local node indices.
    positions: optional list of (x, y) per graph (d3's phyllotaxis start
    otherwise). random_states: optional starting state of each graph's random
    source, d3's lcg (default 1, as every d3 simulation starts)."""

    def __init__(self, graphs, core=None, charge=-30.0, distance=30.0, positions=None, random_states=None):
        self.core = _bind(core or gpu_core())
        c = self.core
        self.h = c.batch_create()
        if not self.h:
            raise RuntimeError("batch_create failed (no GPU?)")
        sizes = np.array([g[0] for g in graphs], dtype=np.int64)
        links = np.array([len(g[1]) for g in graphs], dtype=np.int64)
        self.node_off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
        self.link_off = np.concatenate([[0], np.cumsum(links)]).astype(np.uint32)
        self.graphs = len(graphs)
        check(c.batch_set_graphs(self.h, self.graphs, self.node_off.ctypes.data, self.link_off.ctypes.data))
        self.n = int(self.node_off[-1])
        self.m = int(self.link_off[-1])
        x, y, vx, vy, q = (self._f64(k, self.n) for k in (B_X, B_Y, B_VX, B_VY, B_CHARGE))
        src, tgt = self._u32(B_SOURCE, self.m), self._u32(B_TARGET, self.m)
        ls, ld, lb = (self._f64(k, self.m) for k in (B_LINK_STRENGTH, B_LINK_DISTANCE, B_LINK_BIAS))
        for g, (n, s, t) in enumerate(graphs):
            a, b = int(self.node_off[g]), int(self.node_off[g + 1])
            la, lb_ = int(self.link_off[g]), int(self.link_off[g + 1])
            if positions is not None:
                x[a:b], y[a:b] = positions[g]
            else:
                x[a:b], y[a:b] = phyllotaxis_local(n)
            s = np.asarray(s, dtype=np.int64)
            t = np.asarray(t, dtype=np.int64)
            src[la:lb_] = s + a
            tgt[la:lb_] = t + a
            if len(s):
                ls[la:lb_], lb[la:lb_] = link_defaults(n, s, t)
        vx[:] = 0
        vy[:] = 0
        q[:] = charge
        ld[:] = distance
        if random_states is not None:
            self._states()[:] = random_states
        self.sync_in()

    def _states(self):
        return np.ctypeslib.as_array(ctypes.cast(self.core.batch_ptr(self.h, B_SEED), ctypes.POINTER(ctypes.c_uint64)),
                                     shape=(self.graphs,))

    def random_states(self):
        """Each graph's random state (d3's lcg) after the ticks so far."""
        self.sync_out()
        return self._states().copy()

    def _f64(self, which, count):
        p = self.core.batch_ptr(self.h, which)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_double)), shape=(max(count, 1),))[:count]

    def _u32(self, which, count):
        p = self.core.batch_ptr(self.h, which)
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint32)), shape=(max(count, 1),))[:count]

    def sync_in(self):
        check(self.core.batch_sync(self.h, 0))

    def sync_out(self):
        check(self.core.batch_sync(self.h, 1))

    def scalar(self, which, v=None):
        if v is None:
            return self.core.batch_get_scalar(self.h, which)
        check(self.core.batch_set_scalar(self.h, which, float(v)))
        return self

    def tick(self, k=1):
        check(self.core.batch_tick(self.h, int(k)))
        return self

    def autotune(self, ticks=2):
        """Times `ticks` ticks per candidate launch shape (include/d3fgpu.h,
        batch_autotune) and keeps the fastest; returns the shapes chosen
        (dual-traversal blocks, upward-pass blocks, near-field item blocks) and the
        per-tick milliseconds of the candidates. The ticks (one untimed, then
        `ticks` per candidate) advance the batch."""
        chosen = (ctypes.c_uint32 * 3)()
        ms = (ctypes.c_float * 16)()
        check(self.core.batch_autotune(self.h, ticks, chosen, ms))
        return {"dual_blocks": chosen[0], "level_blocks": chosen[1], "near_extra": chosen[2],
                "candidate_ms": [round(v, 3) for v in ms[:14]]}

    def state(self):
        """x, y, vx, vy of every node as a (4, n) float64 array (a copy)."""
        self.sync_out()
        return np.stack([self._f64(k, self.n).copy() for k in (B_X, B_Y, B_VX, B_VY)])

    def graph_state(self, g, state=None):
        s = self.state() if state is None else state
        return s[:, int(self.node_off[g]):int(self.node_off[g + 1])]

    def stats(self):
        from .core import Stats
        st = Stats()
        check(self.core.batch_stats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self.h:
            self.core.batch_destroy(self.h)
            self.h = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
