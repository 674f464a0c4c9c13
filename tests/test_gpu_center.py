"""forceCenter's sums at sizes where the GPU adds them by binade chunks.

d3 sums x over nodes 0..n-1 in index order (simple.zig:18-22); from 65,536
nodes on, the GPU replaces runs of 32 adds by one exact add whenever the
partial sums provably stay inside one binade, and adds the rest one by one.
numpy's cumsum adds strictly left to right, so its last element is the
reference's sum bit for bit; each case checks x and y after one tick of a
simulation with forceCenter alone (integrate adds the zero velocity, which
turns -0 into +0, as the reference's integrate does). The cases aim at the
chunk rules: rounding ties on every add (the parity of the running sum
decides them), partial sums crossing binades and zero, huge, tiny and
infinite values, all-zero and negative-zero inputs, and sizes around the
65,536 switch and a partial last chunk.
"""

import numpy as np
import pytest

import d3force_amd as d3

pytestmark = pytest.mark.gpu


def expected(x, y, cx=0.0, cy=0.0, strength=1.0):
    n = len(x)
    sx = np.cumsum(x)[-1] if n else 0.0
    sy = np.cumsum(y)[-1] if n else 0.0
    shx = (sx / n - cx) * strength
    shy = (sy / n - cy) * strength
    return (x - shx) + 0.0, (y - shy) + 0.0


def run(core, x, y, cx=0.0, cy=0.0, strength=1.0):
    nodes = d3.Nodes(len(x), x=x.copy(), y=y.copy())
    sim = d3.forceSimulation(nodes, core=core).force("center", d3.forceCenter(cx, cy).strength(strength)).stop()
    sim.tick()
    return nodes.x.copy(), nodes.y.copy()


def case(name, n, seed=0):
    rng = np.random.default_rng(seed)
    if name == "uniform":
        return 1000 * (rng.random(n) - 0.5), 3000 * (rng.random(n) - 0.3)
    if name == "ties":
        # s near 2^40 (ulp 2^-12) plus odd multiples of 2^-13: every add is a tie
        x = (2 * rng.integers(-4000, 4000, n) + 1) * 2.0 ** -13
        x[0] = 2.0 ** 40
        y = (2 * rng.integers(-4000, 4000, n) + 1) * 2.0 ** -13
        y[0] = -(2.0 ** 40) - 2.0 ** 38
        return x, y
    if name == "crossings":
        # a drifting random walk through zero and many binades
        t = np.arange(n)
        x = 5000 * np.sin(t / 700.0) + rng.normal(0, 300, n)
        y = rng.normal(0, 1, n) * np.where(t % 9000 < 4500, 1e4, 1e-3) + 1e-9 * t
        return x, y
    if name == "extremes":
        x = rng.normal(0, 1, n) * 10.0 ** rng.integers(-300, 300, n)
        y = 1000 * (rng.random(n) - 0.5)
        y[n // 3] = np.inf
        return x, y
    if name == "zeros":
        x = np.zeros(n)
        y = -np.zeros(n)
        y[n // 2:] = rng.normal(0, 1e-310, n - n // 2)  # subnormals
        return x, y
    if name == "near power of two":
        x = rng.integers(-3, 4, n) * 2.0 ** -20
        x[0] = 2.0 ** 20 - 2.0 ** -10
        y = rng.integers(-3, 4, n) * 2.0 ** -30
        y[0] = -(2.0 ** 21)
        return x, y
    raise KeyError(name)


CASES = [("uniform", 200000), ("ties", 131072), ("crossings", 250000), ("extremes", 100001),
         ("zeros", 70000), ("near power of two", 65536 + 17), ("uniform", 65535), ("uniform", 65536),
         ("uniform", 1000003)]


@pytest.mark.parametrize("name,n", CASES, ids=["%s-%d" % c for c in CASES])
def test_center_sums_bit_identical(gpu_core, name, n):
    print("center", name, n, flush=True)
    x, y = case(name, n)
    ex, ey = expected(x, y, cx=3.5, cy=-1.25, strength=0.75)
    gx, gy = run(gpu_core, x, y, cx=3.5, cy=-1.25, strength=0.75)
    np.testing.assert_array_equal(gx.view(np.uint64), ex.view(np.uint64))
    np.testing.assert_array_equal(gy.view(np.uint64), ey.view(np.uint64))
