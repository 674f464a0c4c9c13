"""forceSimulation: the host mirror of the reference's src/simulation.js.

The tick loop is d3's (simulation.js:49-89): alpha steps toward alphaTarget,
forces run in insertion order, then integrate. Node arrays are copied into the
core once per ``tick(k)`` call and out once at its end; a force without a
native slot sees the nodes through a copy-out / copy-in around it.
"""

import math
import os

import numpy as np

from . import core as _core
from .lcg import lcg

# Ticks per sim_ticks call (their alphas are computed ahead on the host).
TICK_RUN = 1024

INITIAL_RADIUS = 10.0
INITIAL_ANGLE = math.pi * (3 - math.sqrt(5))  # simulation.js:14-15


def phyllotaxis(n, start=0):
    """d3's initial layout for node i: radius 10 sqrt(0.5 + i), angle
    i * pi (3 - sqrt 5) (simulation.js:96-99), with the C library's cos and
    sin, element by element, so the bits do not depend on numpy's SIMD."""
    x = np.empty(n)
    y = np.empty(n)
    for k in range(n):
        i = start + k
        radius = INITIAL_RADIUS * math.sqrt(0.5 + i)
        angle = i * INITIAL_ANGLE
        x[k] = radius * math.cos(angle)
        y[k] = radius * math.sin(angle)
    return x, y


class Simulation:
    def __init__(self, nodes=None, core=None):
        from .nodes import Nodes
        if nodes is None:
            nodes = Nodes(0)
        self._nodes = nodes
        if core is not None:
            nodes.core = core
        self._alpha = 1.0
        self._alpha_min = 0.001
        self._alpha_decay = 1 - math.pow(self._alpha_min, 1 / 300)
        self._alpha_target = 0.0
        self._velocity_decay = 0.6
        self._forces = {}
        self._random = lcg()
        self._precision = "exact"
        self._fast = None
        # Whole ticks as replays of a captured HIP graph where the core offers
        # it (sim_ticks); D3F_TICK_GRAPH=0 keeps the per-force loop.
        self._captured = os.environ.get("D3F_TICK_GRAPH", "1") != "0"
        self._initialize_nodes()

    # simulation.js:91-105
    def _initialize_nodes(self):
        nd = self._nodes
        n = len(nd)
        if n == 0:
            return
        nd.x[nd.fx_set] = nd.fx[nd.fx_set]
        nd.y[nd.fy_set] = nd.fy[nd.fy_set]
        missing = np.nonzero(np.isnan(nd.x) | np.isnan(nd.y))[0]
        for i in missing:
            radius = INITIAL_RADIUS * math.sqrt(0.5 + int(i))
            angle = int(i) * INITIAL_ANGLE
            nd.x[i] = radius * math.cos(angle)
            nd.y[i] = radius * math.sin(angle)
        still = np.isnan(nd.vx) | np.isnan(nd.vy)
        nd.vx[still] = 0.0
        nd.vy[still] = 0.0

    def _initialize_force(self, force):
        if hasattr(force, "initialize"):
            force.initialize(self._nodes, self._random)
        return force

    # simulation.js:49-89
    def tick(self, iterations=1):
        if not (0 < iterations):
            return self
        ctx = self.begin()
        try:
            self.iterate(iterations, ctx)
        finally:
            self.end(ctx)
        return self

    def begin(self):
        """The start of d3's tick: node arrays copied into the core once."""
        nodes = self._nodes
        ctx = _core.context(nodes)
        ctx.precision = self._fast
        ctx.copy_in(nodes)
        ctx.live = True
        return ctx

    def captured_ticks(self, on=None):
        """Whether tick() runs whole ticks as replays of a captured HIP graph
        (the same bits as the per-force loop) when every force is native."""
        if on is None:
            return self._captured
        self._captured = bool(on)
        return self

    def _native_forces(self, ctx):
        """The forces in order when all run in the core and the random source
        they draw from is the simulation's, else None."""
        forces = list(self._forces.values())
        for f in forces:
            slot = getattr(f, "slot", None)
            if slot is None or slot.ctx is not ctx or not hasattr(f, "prepare"):
                return None
            if getattr(f, "random", self._random) not in (None, self._random):
                return None
        return forces

    def _run_captured(self, iterations, k, ctx):
        """Ticks k.. of the loop through sim_ticks; returns the new k, or None
        when the per-force loop must run them."""
        forces = self._native_forces(ctx)
        if not forces:
            return None
        count = 0
        alphas = []
        a = self._alpha
        while k + count < iterations and count < TICK_RUN:  # simulation.js:63-64, ahead
            count += 1
            a += (self._alpha_target - a) * self._alpha_decay
            alphas.append(a)
        try:
            done = ctx.ticks(forces, alphas, self._velocity_decay, self._random)
        except Exception as e:
            d = getattr(e, "ticks_done", 0)
            if d:
                self._alpha = alphas[d - 1]
            raise
        if done is None:
            return None
        self._alpha = alphas[-1]
        return k + count

    def iterate(self, iterations, ctx):
        """``iterations`` ticks on the core's resident arrays (simulation.js:63-82)."""
        nodes = self._nodes
        k = 0
        captured = self._captured
        while k < iterations:  # simulation.js:63: for (k = 0; k < iterations; ++k)
            if captured:
                nk = self._run_captured(iterations, k, ctx)
                if nk is not None:
                    k = nk
                    continue
                captured = False  # this call's ticks run force by force
            k += 1
            self._alpha += (self._alpha_target - self._alpha) * self._alpha_decay
            for force in list(self._forces.values()):
                slot = getattr(force, "slot", None)
                if slot is not None and slot.ctx is ctx:
                    force(self._alpha)
                    continue
                ctx.live = False
                ctx.copy_out(nodes)
                try:
                    force(self._alpha)
                finally:
                    ctx.copy_in(nodes)
                    ctx.live = True
            ctx.integrate(self._velocity_decay)

    def end(self, ctx):
        """The end of d3's tick: node arrays copied out once."""
        ctx.live = False
        ctx.copy_out(self._nodes)

    def nodes(self, nodes=None):
        if nodes is None:
            return self._nodes
        if nodes.core is None:
            nodes.core = self._nodes.core
        self._nodes = nodes
        self._initialize_nodes()
        for f in self._forces.values():
            self._initialize_force(f)
        return self

    def alpha(self, v=None):
        if v is None:
            return self._alpha
        self._alpha = float(v)
        return self

    def alpha_min(self, v=None):
        if v is None:
            return self._alpha_min
        self._alpha_min = float(v)
        return self

    def alpha_decay(self, v=None):
        if v is None:
            return self._alpha_decay
        self._alpha_decay = float(v)
        return self

    def alpha_target(self, v=None):
        if v is None:
            return self._alpha_target
        self._alpha_target = float(v)
        return self

    def velocity_decay(self, v=None):  # stores the multiplier (simulation.js:146)
        if v is None:
            return 1 - self._velocity_decay
        self._velocity_decay = 1 - float(v)
        return self

    def random_source(self, fn=None):
        if fn is None:
            return self._random
        self._random = fn
        for f in self._forces.values():
            self._initialize_force(f)
        return self

    def force(self, name, f=None, remove=False):
        if remove:
            self._forces.pop(name, None)
            return self
        if f is None:
            return self._forces.get(name)
        self._forces[name] = self._initialize_force(f)
        return self

    def stop(self):
        return self

    # simulation.js:187-200
    def precision(self, mode=None, options=None):
        if mode is None:
            return self._precision
        if mode not in ("exact", "fast"):
            raise ValueError('precision must be "exact" or "fast"')
        self._precision = mode
        if mode == "fast":
            o = options or {}
            order = o.get("order") or 8
            opening = o.get("opening") or 1.5
            order = float(order)
            self._fast = {
                "order": min(24, max(1, _js_round(order if order == order else 8))),
                "opening": max(1.01, float(opening) if float(opening) == float(opening) else 1.5),
            }
        else:
            self._fast = None
        _core.context(self._nodes).precision = self._fast
        return self


def _js_round(v):
    """JS Math.round: halves round toward +infinity."""
    return float(math.floor(v + 0.5))


def forceSimulation(nodes=None, core=None):
    return Simulation(nodes, core)
