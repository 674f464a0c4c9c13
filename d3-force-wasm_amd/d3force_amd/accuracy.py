"""Force error of fast mode against the FP64 direct sum.

The measurement of the reference's test/fast/error-test.js:33-70 at sizes a
host loop cannot reach: one tick of forceManyBody alone with velocityDecay 0
from v = 0 on a copy of the layout (so v after it is the force), against the
library's FP64 direct sum (``force_direct_sum``) over a fixed sample of
nodes, as the relative RMS sqrt(sum |v - f|^2 / sum |f|^2) of
error-test.js:60-70 (fmm.zig:981-994 computes the same quantity).
"""

import ctypes

import numpy as np

from .core import check, gpu_core
from .forces import forceManyBody
from .nodes import Nodes
from .simulation import forceSimulation


def fast_force_error(x, y, sample_size=10000, seed=3, order=8, opening=1.5, core=None):
    """Relative RMS error of the fast many-body force on the layout (x, y)
    over ``sample_size`` nodes drawn without replacement with ``seed``.
    Returns (error, sample size)."""
    core = core or gpu_core()
    n = len(x)
    probe = Nodes(n, x=np.array(x, dtype=np.float64), y=np.array(y, dtype=np.float64),
                  vx=np.zeros(n), vy=np.zeros(n))
    charge = forceManyBody()
    sim = forceSimulation(probe, core=core).velocity_decay(0).force("charge", charge).stop()
    sim.precision("fast", {"order": order, "opening": opening})
    sample = np.sort(np.random.default_rng(seed).choice(n, min(n, sample_size), replace=False)).astype(np.uint32)
    ref = np.empty(2 * len(sample))
    ctx = sim.begin()
    try:
        charge.push_scalars()
        check(core.force_direct_sum(ctx.h, charge.slot.fh, 1.0, sample.ctypes.data_as(ctypes.c_void_p),
                                    len(sample), ref.ctypes.data_as(ctypes.c_void_p)))
        sim.iterate(1, ctx)
    finally:
        sim.end(ctx)
    ref *= sim.alpha()  # the sum is linear in alpha; the tick used this alpha
    fx, fy = ref[0::2], ref[1::2]
    ex, ey = probe.vx[sample] - fx, probe.vy[sample] - fy
    return float(np.sqrt(np.sum(ex * ex + ey * ey) / np.sum(fx * fx + fy * fy))), len(sample)
