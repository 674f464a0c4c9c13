"""Loader and sync layer: the host-side mirror of the reference's src/core.js.

A *core* is a shared library exporting the tick engine's C ABI
(include/d3fgpu.h): the MI355X library ``libd3fgpu.so`` by default. Each node
table gets one ``Context`` (core.js:89-164) holding the core's simulation
handle and numpy views of its x, y, vx, vy, fx, fy, fixed arrays; each native
force holds a ``Slot`` (core.js:174-232) with its force handle.

Differences from core.js, all forced by the GPU boundary:
  * ``copy_in`` ends with ``sim_sync(h, SYNC_H2D)`` and ``copy_out`` starts with
    ``sim_sync(h, SYNC_D2H)``: typed-array writes cannot be intercepted, so the
    host says when the host arrays are authoritative.
  * the random source is a C callback (``core_set_random``) in place of the
    wasm import ``env.random`` (core.js:11-14); ``Slot.run`` swaps it per run
    exactly as core.js:223-229 swaps ``random``.
"""

import ctypes
import os
import weakref

import numpy as np

# Force kinds, scalar slots and array slots (src/core/force.zig:8, :13, :21;
# core.js:33-37).
CENTER, COLLIDE, LINK, MANY_BODY, RADIAL, X, Y = range(7)
(STRENGTH, CX, CY, THETA2, DISTANCE_MIN2, DISTANCE_MAX2, ITERATIONS,
 PRECISION, ORDER, OPENING, SEED) = range(11)
F0, F1, F2, SOURCE, TARGET = range(5)
SCALARS = 11
# Float parameter arrays per force kind (force.zig:54-61, core.js:40).
FLOATS = [0, 1, 3, 1, 2, 2, 2]

OK, GROWN = 0, 1
NOT_CAPTURED = 2  # sim_ticks: nothing ran, use the per-force loop
BAD_HANDLE, OUT_OF_MEMORY, BAD_ARGUMENT, WORKER_FAILED = -1, -2, -3, -4
_STATUSES = {-1: "BAD_HANDLE", -2: "OUT_OF_MEMORY", -3: "BAD_ARGUMENT", -4: "WORKER_FAILED"}

SYNC_H2D, SYNC_D2H = 0, 1

# sim_ptr slots (src/core/sim.zig:14).
P_X, P_Y, P_VX, P_VY, P_FX, P_FY, P_FIXED, P_PERM, P_RNG, P_SNAP = range(10)


class CoreError(RuntimeError):
    """A negative core status, named as core.js:42-48 names it."""

    def __init__(self, status):
        self.status = status
        super().__init__(_STATUSES.get(status, "core status %d" % status))


def check(status):
    """core.js:45-48: throw on a negative status, else pass it through. A GPU
    failure carries the library's account of its cause (core_last_error)."""
    if status < 0:
        err = CoreError(status)
        if status in (OUT_OF_MEMORY, WORKER_FAILED) and _gpu_core is not None and _gpu_core.last_error:
            cause = _gpu_core.last_error()
            if cause:
                err.args = ("%s: %s" % (err.args[0], cause.decode(errors="replace")),)
        raise err
    return status


_RANDOM_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p)

PHASES = ("tree", "accumulate", "traversal", "link", "center", "integrate", "center_sum", "fmm", "collide")


class Stats(ctypes.Structure):
    """d3f_stats of include/d3fgpu.h: per-phase kernel time and launches since
    the last call, and the Barnes-Hut traversal's work counters."""
    _fields_ = [("ms", ctypes.c_double * len(PHASES)), ("launches", ctypes.c_uint64 * len(PHASES)),
                ("visits", ctypes.c_uint64), ("accepts", ctypes.c_uint64), ("pairs", ctypes.c_uint64),
                ("flagged", ctypes.c_uint64), ("link_levels", ctypes.c_uint64),
                ("m2l", ctypes.c_uint64), ("p2p_pairs", ctypes.c_uint64), ("fmm_cells", ctypes.c_uint64),
                ("collide_serial", ctypes.c_uint64), ("collide_status", ctypes.c_uint64)]

    def as_dict(self):
        return {"ms": {p: self.ms[k] for k, p in enumerate(PHASES)},
                "launches": {p: int(self.launches[k]) for k, p in enumerate(PHASES)},
                "visits": int(self.visits), "accepts": int(self.accepts), "pairs": int(self.pairs),
                "flagged": int(self.flagged), "link_levels": int(self.link_levels),
                "m2l": int(self.m2l), "p2p_pairs": int(self.p2p_pairs), "fmm_cells": int(self.fmm_cells),
                "collide_serial": int(self.collide_serial), "collide_status": int(self.collide_status)}

_SIGNATURES = {
    "core_set_random": (None, [_RANDOM_FN, ctypes.c_void_p]),
    "sim_create": (ctypes.c_uint32, []),
    "sim_destroy": (ctypes.c_int32, [ctypes.c_uint32]),
    "sim_set_node_count": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "sim_ptr": (ctypes.c_void_p, [ctypes.c_uint32, ctypes.c_uint32]),
    "sim_sync": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "sim_snapshot": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "sim_integrate": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_double]),
    "tree_build": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "tree_info": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32]),
    "tree_order": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]),
    "force_create": (ctypes.c_uint32, [ctypes.c_uint32, ctypes.c_uint32]),
    "force_destroy": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32]),
    "force_set_count": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "force_ptr": (ctypes.c_void_p, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "force_set_scalar": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double]),
    "force_get_scalar": (ctypes.c_double, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
    "force_run": (ctypes.c_int32, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double]),
}

_HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIBRARY = os.path.join(os.path.dirname(_HERE), "libd3fgpu.so")


class Core:
    """One loaded core library and its exported ABI.

    ``prefix`` is prepended to every export name, so a library that exports
    the same ABI under another prefix can be driven by the same facade.
    """

    def __init__(self, path, prefix=""):
        if not os.path.exists(path):
            raise OSError("core library not built: %s (run __graft_entry__.build())" % path)
        self.path = path
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        self.random = None  # the running force's random source (core.js:12)
        self._callback = _RANDOM_FN(self._draw)
        self.core_set_random(self._callback, None)
        self.contexts = weakref.WeakKeyDictionary()
        if hasattr(self.lib, prefix + "sim_stats"):
            self.sim_stats = self.bind("sim_stats", ctypes.c_int32, [ctypes.c_uint32, ctypes.POINTER(Stats)])
            self.core_set_timing = self.bind("core_set_timing", None, [ctypes.c_int])
            self.core_synchronize = self.bind("core_synchronize", ctypes.c_int32, [])
        self.last_error = None
        if hasattr(self.lib, prefix + "core_last_error"):
            self.last_error = self.bind("core_last_error", ctypes.c_char_p, [])
        # A core whose steps can fail (the GPU library) snapshots each step's
        # inputs so the facade can rerun it (pool.js:73-84).
        self.can_fail = hasattr(self.lib, prefix + "core_fail_next")
        if self.can_fail:
            self.core_fail_next = self.bind("core_fail_next", ctypes.c_int32,
                                            [ctypes.c_int32, ctypes.c_uint32, ctypes.c_uint32])
        # Whole ticks captured as a HIP graph (a GPU-library extension).
        self.sim_ticks = None
        if hasattr(self.lib, prefix + "sim_ticks"):
            self.sim_ticks = self.bind("sim_ticks", ctypes.c_int32,
                                       [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_uint32)])
            self.sim_tick_counters = self.bind("sim_tick_counters", ctypes.c_int32,
                                               [ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)])
        if hasattr(self.lib, prefix + "force_direct_sum"):
            self.force_direct_sum = self.bind("force_direct_sum", ctypes.c_int32,
                                              [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_void_p,
                                               ctypes.c_uint32, ctypes.c_void_p])

    def stats(self, h):
        """Per-phase kernel times and work counters since the last call."""
        st = Stats()
        check(self.sim_stats(h, ctypes.byref(st)))
        return st.as_dict()

    def _draw(self, _ctx):
        return float(self.random())

    def bind(self, name, res, args):
        """Binds an extra export (library-specific extensions)."""
        fn = getattr(self.lib, self.prefix + name)
        fn.restype = res
        fn.argtypes = args
        return fn


_gpu_core = None


def gpu_core():
    """The MI355X core, loaded once. Raises when the library is missing: there
    is no CPU fallback on the product path."""
    global _gpu_core
    if _gpu_core is None:
        _gpu_core = Core(GPU_LIBRARY)
    return _gpu_core


def _view(ptr, dtype, n):
    if n == 0 or not ptr:
        return np.zeros(0, dtype=dtype)
    ctype = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
    buf = (ctype * n).from_address(ptr)
    return np.ctypeslib.as_array(buf)


class Context:
    """One simulation region in a core for one node table (core.js:89-164)."""

    def __init__(self, core):
        self.core = core
        self.h = core.sim_create()
        if not self.h:
            raise CoreError(OUT_OF_MEMORY)
        self.n = 0
        self.live = False  # true while a tick owns the arrays
        self.ptrs = None
        self.precision = None  # None = exact; dict(order, opening) = fast
        self.slots = []
        self.failover = True  # pool.js:73-84: rerun a failed step (on cores whose steps can fail)
        self.failovers = 0
        self._finalizer = weakref.finalize(self, core.sim_destroy, self.h)

    def resize(self, n):  # core.js:102-104
        if n != self.n:
            check(self.core.sim_set_node_count(self.h, n))
            self.n = n
            self.ptrs = None

    def views(self):  # core.js:106-119, refreshed when an array moved
        core, h, n = self.core, self.h, self.n
        ptrs = tuple(core.sim_ptr(h, w) for w in range(7))
        if ptrs == self.ptrs:
            return
        self.ptrs = ptrs
        self.x = _view(ptrs[0], np.float64, n)
        self.y = _view(ptrs[1], np.float64, n)
        self.vx = _view(ptrs[2], np.float64, n)
        self.vy = _view(ptrs[3], np.float64, n)
        self.fx = _view(ptrs[4], np.float64, n)
        self.fy = _view(ptrs[5], np.float64, n)
        self.fixed = _view(ptrs[6], np.uint8, n)

    def copy_in(self, nodes):  # core.js:121-131
        n = len(nodes)
        self.resize(n)
        self.views()
        if n:
            self.x[:] = nodes.x
            self.y[:] = nodes.y
            self.vx[:] = nodes.vx
            self.vy[:] = nodes.vy
            self.fx[:] = nodes.fx
            self.fy[:] = nodes.fy
            self.fixed[:] = nodes.fx_set.astype(np.uint8) | (nodes.fy_set.astype(np.uint8) << 1)
        check(self.core.sim_sync(self.h, SYNC_H2D))

    def copy_out(self, nodes):  # core.js:133-140
        check(self.core.sim_sync(self.h, SYNC_D2H))
        self.views()
        if self.n:
            nodes.x[:] = self.x
            nodes.y[:] = self.y
            nodes.vx[:] = self.vx
            nodes.vy[:] = self.vy

    def step(self, slot, arg):
        """core.js:155-159 with pool.js:73-84's failover: on a core whose steps
        can fail, the step's vx, vy (and x, y for integrate) are snapshotted
        first; if the step returns WORKER_FAILED or OUT_OF_MEMORY, the
        simulation moves to a fresh simulation of the same core with the
        snapshot restored and the step runs again there, so the tick still
        returns d3's bits. A second failure raises, naming its cause."""
        core = self.core
        guard = self.failover and getattr(core, "can_fail", False)
        if guard:
            check(core.sim_snapshot(self.h, 0 if slot is not None else 1))
        s = self._run(slot, arg)
        if guard and s in (WORKER_FAILED, OUT_OF_MEMORY):
            self.failovers += 1
            self.move(restore=1 if slot is not None else 2)
            s = self._run(slot, arg)
        check(s)

    def _run(self, slot, arg):
        if slot is not None:
            return self.core.force_run(self.h, slot.fh, arg)
        return self.core.sim_integrate(self.h, arg)

    def move(self, restore=0):
        """pool.js:93-150 moveTo: a fresh simulation in the same core takes
        over this one's node arrays and every bound force's arrays and
        scalars verbatim; `restore` takes vx, vy (and x, y when 2) from the
        snapshot of the failed step."""
        core, old, n = self.core, self.h, self.n
        snap = None
        if restore:
            snap = _view(core.sim_ptr(old, P_SNAP), np.float64, (4 if restore == 2 else 2) * n).copy()
        if restore != 2:
            check(core.sim_sync(old, SYNC_D2H))  # the step's x, y (integrate restores them too)
        self.views()
        x, y, vx, vy = self.x.copy(), self.y.copy(), self.vx.copy(), self.vy.copy()
        fx, fy, fixed = self.fx.copy(), self.fy.copy(), self.fixed.copy()
        if snap is not None and n:
            vx, vy = snap[:n], snap[n:2 * n]
            if restore == 2:
                x, y = snap[2 * n:3 * n], snap[3 * n:4 * n]
        h = core.sim_create()
        if not h:
            raise CoreError(OUT_OF_MEMORY)
        handles = []
        try:
            check(core.sim_set_node_count(h, n))
            for slot in self.slots:
                fh = core.force_create(h, slot.kind)
                if not fh:
                    raise CoreError(OUT_OF_MEMORY)
                m = slot.m if slot.kind == LINK else n
                if slot.kind == LINK:
                    check(core.force_set_count(h, fh, m))
                for which in range(FLOATS[slot.kind]):
                    _view(core.force_ptr(h, fh, which), np.float64, m)[:] = \
                        _view(core.force_ptr(old, slot.fh, which), np.float64, m)
                if slot.kind == LINK:
                    for which in (SOURCE, TARGET):
                        _view(core.force_ptr(h, fh, which), np.uint32, m)[:] = \
                            _view(core.force_ptr(old, slot.fh, which), np.uint32, m)
                for which in range(SCALARS):
                    core.force_set_scalar(h, fh, which, core.force_get_scalar(old, slot.fh, which))
                handles.append(fh)
        except Exception:
            core.sim_destroy(h)
            raise
        for slot, fh in zip(self.slots, handles):
            slot.fh = fh
        self._finalizer()  # destroys the old simulation
        self.h = h
        self._finalizer = weakref.finalize(self, core.sim_destroy, h)
        self.ptrs = None
        self.views()
        if n:
            self.x[:], self.y[:], self.vx[:], self.vy[:] = x, y, vx, vy
            self.fx[:], self.fy[:], self.fixed[:] = fx, fy, fixed
        check(core.sim_sync(h, SYNC_H2D))

    def integrate(self, velocity_decay):  # core.js:161-163
        self.step(None, velocity_decay)

    def ticks(self, forces, alphas, velocity_decay, rand):
        """Whole ticks in one library call (sim_ticks): for each alpha, every
        force of `forces` (native slots of this context, in order), then
        integrate, as the per-force loop runs them, bit for bit and draw for
        draw of `rand`. Returns the ticks run, or None when the library asks
        for the per-force loop (nothing ran then)."""
        core = self.core
        if core.sim_ticks is None:
            return None
        for f in forces:
            f.prepare()
            f.slot.prepare(self.precision)
        fhs = (ctypes.c_uint32 * len(forces))(*[f.slot.fh for f in forces])
        a = np.ascontiguousarray(alphas, dtype=np.float64)
        done = ctypes.c_uint32(0)
        previous = core.random
        core.random = rand
        try:
            s = core.sim_ticks(self.h, fhs, len(forces), len(a), a.ctypes.data, velocity_decay, ctypes.byref(done))
        finally:
            core.random = previous
        if s == NOT_CAPTURED:
            return None
        if s < 0:
            err = None
            try:
                check(s)
            except CoreError as e:
                err = e
            err.ticks_done = int(done.value)
            raise err
        return int(done.value)

    def tick_counters(self):
        """(captures, replayed ticks, ticks rolled back and run eagerly)."""
        out = (ctypes.c_uint64 * 3)()
        check(self.core.sim_tick_counters(self.h, out))
        return tuple(int(v) for v in out)


def context(nodes):
    """core.js:166-170: the node table's Context in the table's core."""
    core = nodes.core
    if core is None:
        core = nodes.core = gpu_core()
    c = core.contexts.get(nodes)
    if c is None:
        c = Context(core)
        core.contexts[nodes] = c
    return c


class Slot:
    """A native force's handle in the core (core.js:174-232)."""

    def __init__(self, kind):
        self.kind = kind
        self.ctx = None
        self.fh = 0
        self.m = 0

    def bind(self, nodes):  # core.js:182-191
        ctx = context(nodes)
        if ctx is not self.ctx:
            if self.ctx is not None:
                self.ctx.core.force_destroy(self.ctx.h, self.fh)
                if self in self.ctx.slots:
                    self.ctx.slots.remove(self)
            self.fh = ctx.core.force_create(ctx.h, self.kind)
            if not self.fh:
                raise CoreError(OUT_OF_MEMORY)
            self.ctx, self.m = ctx, 0
            ctx.slots.append(self)
        ctx.resize(len(nodes))

    def count(self, m):  # core.js:193-196
        check(self.ctx.core.force_set_count(self.ctx.h, self.fh, m))
        self.m = m

    def f64(self, which, m):  # core.js:199-202
        return _view(self.ctx.core.force_ptr(self.ctx.h, self.fh, which), np.float64, m)

    def u32(self, which, m):  # core.js:204-207
        return _view(self.ctx.core.force_ptr(self.ctx.h, self.fh, which), np.uint32, m)

    def scalar(self, which, v):  # core.js:209-211
        self.ctx.core.force_set_scalar(self.ctx.h, self.fh, which, float(v))

    def prepare(self, fast):
        """The precision scalars run() sets before a run (core.js:216-221)."""
        self.scalar(PRECISION, 1 if fast else 0)
        if fast:
            self.scalar(ORDER, fast["order"])
            self.scalar(OPENING, fast["opening"])

    def run(self, nodes, alpha, rand=None):  # core.js:213-231
        ctx = self.ctx
        fast = ctx.precision
        if not ctx.live:
            ctx.copy_in(nodes)
        self.scalar(PRECISION, 1 if fast else 0)
        if fast:
            self.scalar(ORDER, fast["order"])
            self.scalar(OPENING, fast["opening"])
            if self.kind in (MANY_BODY, LINK, COLLIDE):
                self.scalar(SEED, rand())
        core = ctx.core
        previous = core.random
        core.random = rand
        try:
            ctx.step(self, alpha)
        finally:
            core.random = previous
            if not ctx.live:
                ctx.copy_out(nodes)
