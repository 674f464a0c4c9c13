"""Known answers and mathematical identities that pin the exact-mode tick
(tree charges, Barnes-Hut traversal, jiggle draws, Gauss-Seidel link,
center, integrate, array semantics) independently of any restatement of the
reference's code. Each check takes a core (the oracle, or the MI355X
library) and asserts against values fixed by the reference's own tests or by
the mathematics:

* src/core/tests.zig:128-161  array semantics (NaN padding, links not
  node-sized, BAD_ARGUMENT on out-of-range endpoints);
* src/core/tests.zig:163-182  integrate known answer;
* src/core/tests.zig:190-239  fast link == exact link bit for bit on a
  perfect matching (the fast link is pinned by the reference's fast-mode
  digests, so this pins the exact link's per-edge arithmetic);
* src/core/tests.zig:241-284  fast collide == exact collide on isolated
  pairs, every node moved;
* test/d3/center-test.js, test/d3/x-test.js, test/d3/simulation-test.js:10-17
  (phyllotaxis) known answers;
* Barnes-Hut identities: theta = inf sees only the root (closed form from the
  weighted centroid), theta = 0 is the O(n^2) direct sum, the accumulated
  charge of every cell is the sum of its points' strengths at their weighted
  centroid, the three-node exact tie of test/parity/scenarios.js:176-185 and
  two coincident nodes whose jiggle draws follow d3's order, both derived by
  hand;
* Gauss-Seidel link identities: one iteration at alpha = strength = 1 puts
  an isolated edge at exactly its distance and keeps the bias-weighted point
  of its shifted endpoints; on a path the last edge is exact;
* test/fast/error-test.js:60-81: fast mode's RMS error against the direct
  sum is at most 1e-3 and at most Barnes-Hut's own.
"""

import math

import numpy as np

import d3force_amd as d3
from d3force_amd import core as C
from d3force_amd.core import _view
from d3force_amd.lcg import lcg

NAN = float("nan")
INF = float("inf")


class Raw:
    """One simulation driven through the bare ABI (as the reference's
    src/core/tests.zig drives sim.zig and the force files)."""

    def __init__(self, core, n):
        self.core = core
        self.h = core.sim_create()
        assert self.h
        self.n = 0
        self.set_n(n)

    def set_n(self, n):
        st = self.core.sim_set_node_count(self.h, n)
        self.n = n
        return st

    def arr(self, which):
        dt = np.uint8 if which == C.P_FIXED else np.float64
        return _view(self.core.sim_ptr(self.h, which), dt, self.n)

    def put(self, x, y, vx=None, vy=None, fx=None, fy=None, fixed=None):
        n = self.n
        self.arr(C.P_X)[:] = x
        self.arr(C.P_Y)[:] = y
        self.arr(C.P_VX)[:] = 0.0 if vx is None else vx
        self.arr(C.P_VY)[:] = 0.0 if vy is None else vy
        self.arr(C.P_FX)[:] = NAN if fx is None else fx
        self.arr(C.P_FY)[:] = NAN if fy is None else fy
        self.arr(C.P_FIXED)[:] = np.zeros(n, np.uint8) if fixed is None else fixed
        assert self.core.sim_sync(self.h, C.SYNC_H2D) == 0

    def get(self):
        assert self.core.sim_sync(self.h, C.SYNC_D2H) == 0
        return [self.arr(w).copy() for w in (C.P_X, C.P_Y, C.P_VX, C.P_VY)]

    def force(self, kind, count=None):
        fh = self.core.force_create(self.h, kind)
        assert fh
        if count is not None:
            assert self.core.force_set_count(self.h, fh, count) == C.GROWN
        return fh

    def f64(self, fh, which, m):
        return _view(self.core.force_ptr(self.h, fh, which), np.float64, m)

    def u32(self, fh, which, m):
        return _view(self.core.force_ptr(self.h, fh, which), np.uint32, m)

    def scalar(self, fh, which, v):
        assert self.core.force_set_scalar(self.h, fh, which, float(v)) == 0

    def run(self, fh, alpha, random=None):
        self.core.random = random or (lambda: 0.5)
        try:
            return self.core.force_run(self.h, fh, alpha)
        finally:
            self.core.random = None

    def close(self):
        self.core.sim_destroy(self.h)


def bitwise(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return bool(np.all((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))))


# ---------------------------------------------------------------- tests.zig


def check_node_sized_arrays_follow_node_count(core):
    """tests.zig:128-145."""
    s = Raw(core, 0)
    fh = s.force(C.X, 2)
    s.f64(fh, C.F0, 2)[0] = 1
    s.f64(fh, C.F1, 2)[1] = 2
    assert s.set_n(5) == C.GROWN
    f0, f1 = s.f64(fh, C.F0, 5), s.f64(fh, C.F1, 5)
    assert f0[0] == 1 and f1[1] == 2
    assert np.isnan(f0[2:]).all() and np.isnan(f1[2:]).all()
    # Shrinking keeps the arrays (count stays 5); growing back to 5 then
    # needs no new slots, so it reports OK.
    assert s.set_n(3) == C.OK
    assert s.set_n(5) == C.OK
    f0 = s.f64(fh, C.F0, 5)
    assert f0[0] == 1 and np.isnan(f0[2:]).all()
    s.close()


def check_links_not_node_sized_and_bad_index(core):
    """tests.zig:147-161: a link force keeps its count when n changes, and an
    endpoint >= n makes force_run return BAD_ARGUMENT."""
    s = Raw(core, 3)
    fh = s.force(C.LINK, 1)
    s.set_n(10)
    s.put(np.arange(10.0), np.zeros(10))
    src, tgt = s.u32(fh, C.SOURCE, 1), s.u32(fh, C.TARGET, 1)
    src[0], tgt[0] = 0, 10
    for w, v in ((C.F0, 1.0), (C.F1, 30.0), (C.F2, 0.5)):
        s.f64(fh, w, 1)[0] = v
    s.scalar(fh, C.ITERATIONS, 1)

    def never():
        raise AssertionError("BAD_ARGUMENT must come before any draw")
    assert s.run(fh, 1.0, never) == C.BAD_ARGUMENT
    s.close()


def check_integrate_fixed_rule(core):
    """tests.zig:163-182: integrate applies the fixed-position rule per axis."""
    s = Raw(core, 2)
    s.put(x=[1, 1], y=[2, 2], vx=[10, 10], vy=[20, 20], fx=[-5, -5], fy=[-6, -6],
          fixed=np.array([0, 2], np.uint8))
    assert core.sim_integrate(s.h, 0.5) == 0
    x, y, vx, vy = s.get()
    assert list(x) == [6, 6] and list(vx) == [5, 5]
    assert list(y) == [12, -6] and list(vy) == [10, 0]
    s.close()


def _matching_layout(n):
    r = lcg()
    x, y, vx, vy = (np.empty(n) for _ in range(4))
    for i in range(n):
        x[i] = 500 * (r() - 0.5)
        y[i] = 500 * (r() - 0.5)
        vx[i] = r() - 0.5
        vy[i] = r() - 0.5
    x[7], y[7] = x[6], y[6]  # a zero-length edge, so the exact link jiggles
    return x, y, vx, vy


def check_fast_link_equals_exact_on_matching(core):
    """tests.zig:190-239: with one edge per node Jacobi and Gauss-Seidel see
    the same operands, so the fast link sweep (n >= 2000) and the exact
    link agree bit for bit, except on the jiggled zero-length edge."""
    n = 2400
    x, y, vx, vy = _matching_layout(n)
    out = []
    for precision in (1, 0):
        s = Raw(core, n)
        s.put(x, y, vx, vy)
        fh = s.force(C.LINK, n // 2)
        s.u32(fh, C.SOURCE, n // 2)[:] = np.arange(0, n, 2)
        s.u32(fh, C.TARGET, n // 2)[:] = np.arange(1, n, 2)
        s.f64(fh, C.F0, n // 2)[:] = 0.7
        s.f64(fh, C.F1, n // 2)[:] = 30
        s.f64(fh, C.F2, n // 2)[:] = 0.5
        s.scalar(fh, C.ITERATIONS, 1)
        s.scalar(fh, C.SEED, 0.25)
        s.scalar(fh, C.PRECISION, precision)
        assert s.run(fh, 0.3, lambda: 0.25) == 0
        out.append(s.get())
        s.close()
    keep = np.ones(n, bool)
    keep[[6, 7]] = False
    for k in (2, 3):
        assert bitwise(out[0][k][keep], out[1][k][keep])
    # The jiggled edge moved in both (by different draws).
    assert out[0][2][6] != vx[6] and out[1][2][6] != vx[6]


def check_fast_collide_matches_exact_on_isolated_pairs(core):
    """tests.zig:241-284: pairs 10 apart on a 100-unit grid, radius 6, so
    each pair overlaps and nothing else is in reach; the grid collide
    (fast) and the tree collide (exact) agree to 1e-12 and move every node."""
    n = 2000
    i = np.arange(n)
    pair = (i // 2).astype(np.float64)
    x = np.mod(pair, 40.0) * 100 + np.where(i % 2 == 1, 10.0, 0.0)
    y = np.floor(pair / 40.0) * 100 + np.where(i % 2 == 1, 3.0, 0.0)
    out = []
    for precision in (1, 0):
        s = Raw(core, n)
        s.put(x, y)
        fh = s.force(C.COLLIDE, n)
        s.f64(fh, C.F0, n)[:] = 6
        s.scalar(fh, C.STRENGTH, 1)
        s.scalar(fh, C.ITERATIONS, 1)
        s.scalar(fh, C.SEED, 0.25)
        s.scalar(fh, C.PRECISION, precision)
        assert s.run(fh, 0.0, lambda: 0.25) == 0
        out.append(s.get())
        s.close()
    for k in (2, 3):
        np.testing.assert_allclose(out[1][k], out[0][k], rtol=1e-12, atol=0)
    assert int(np.count_nonzero(out[1][2])) == n


# ---------------------------------------------------------------- test/d3


def _nodes(dicts, core):
    t = d3.Nodes.from_dicts(dicts)
    t.core = core
    return t


def _node_equal(t, i, want, delta=1e-6):  # test/d3/asserts.js:7-16
    got = {"x": t.x[i], "y": t.y[i], "vx": t.vx[i], "vy": t.vy[i]}
    for k, v in want.items():
        if k in got:
            assert abs(got[k] - v) < delta, "node %d %s: %r vs %r" % (i, k, got[k], v)
        elif k == "fx":
            assert t.fx_set[i] and abs(t.fx[i] - v) < delta
        elif k == "fy":
            assert t.fy_set[i] and abs(t.fy[i] - v) < delta


def check_center_known_answers(core):
    """test/d3/center-test.js:4-22."""
    t = _nodes([{"x": 100.0, "y": 0.0}, {"x": 200.0, "y": 0.0}, {"x": 300.0, "y": 0.0}], core)
    d3.forceSimulation(t).force("center", d3.forceCenter(0, 0)).tick()
    _node_equal(t, 0, {"x": -100, "y": 0, "vx": 0, "vy": 0})
    _node_equal(t, 1, {"x": 0, "y": 0, "vx": 0, "vy": 0})
    _node_equal(t, 2, {"x": 100, "y": 0, "vx": 0, "vy": 0})
    t = _nodes([{"fx": 0.0, "fy": 0.0}, {}, {}], core)
    d3.forceSimulation(t).force("center", d3.forceCenter()).tick()
    _node_equal(t, 0, {"fx": 0, "fy": 0, "x": 0, "y": 0, "vx": 0, "vy": 0})


def check_axis_known_answers(core):
    """test/d3/x-test.js: forceX / forceY pull toward their target over 30
    ticks, a node on the target stays exactly there, fixed nodes stay put,
    and a per-node target accessor is honoured."""
    for axis, other in (("x", "y"), ("y", "x")):
        t = _nodes([{axis: 100.0, other: 0.0}, {axis: 200.0, other: 0.0}, {axis: 300.0, other: 0.0}], core)
        f = d3.forceX(200) if axis == "x" else d3.forceY(200)
        d3.forceSimulation(t).force(axis, f).tick(30)
        p, v = getattr(t, axis), getattr(t, "v" + axis)
        assert p[0] > 190 and v[0] > 0
        assert p[1] == 200 and v[1] == 0
        assert p[2] < 210 and v[2] < 0
        # accessor: targets 300, 200, 100 (x-test.js:56-82)
        t = _nodes([{axis: 100.0, other: 0.0}, {axis: 200.0, other: 0.0}, {axis: 300.0, other: 0.0}], core)
        f = d3.forceX([300.0, 200.0, 100.0]) if axis == "x" else d3.forceY([300.0, 200.0, 100.0])
        d3.forceSimulation(t).force(axis, f).tick(30)
        p, v = getattr(t, axis), getattr(t, "v" + axis)
        assert p[0] > 290 and v[0] > 0
        assert p[1] == 200 and v[1] == 0
        assert p[2] < 110 and v[2] < 0
    # fixed positions (x-test.js:34-50)
    for make in (d3.forceX, d3.forceY):
        t = _nodes([{"fx": 0.0, "fy": 0.0}, {}, {}], core)
        d3.forceSimulation(t).force("f", make(200)).tick()
        _node_equal(t, 0, {"fx": 0, "fy": 0, "x": 0, "y": 0, "vx": 0, "vy": 0})


def check_phyllotaxis_known_answer(core):
    """test/d3/simulation-test.js:10-17: indices and phyllotaxis positions,
    zero speed."""
    t = _nodes([{}, {}, {}], core)
    d3.forceSimulation(t)
    _node_equal(t, 0, {"x": 7.0710678118654755, "y": 0, "vx": 0, "vy": 0})
    _node_equal(t, 1, {"x": -9.03088751750192, "y": 8.27303273571596, "vx": 0, "vy": 0})
    _node_equal(t, 2, {"x": 1.3823220809823638, "y": -15.750847141167634, "vx": 0, "vy": 0})


# ---------------------------------------------------------------- Barnes-Hut


def _layout(n, seed, spread=500.0):
    r = lcg(seed)
    x = np.array([spread * (r() - 0.5) for _ in range(n)])
    y = np.array([spread * (r() - 0.5) for _ in range(n)])
    return x, y


def _many_body(core, x, y, strengths, theta2, dmin2, dmax2, alpha, random=None):
    n = len(x)
    s = Raw(core, n)
    s.put(x, y)
    fh = s.force(C.MANY_BODY, n)
    s.f64(fh, C.F0, n)[:] = strengths
    s.scalar(fh, C.THETA2, theta2)
    s.scalar(fh, C.DISTANCE_MIN2, dmin2)
    s.scalar(fh, C.DISTANCE_MAX2, dmax2)
    assert s.run(fh, alpha, random) == 0
    _, _, vx, vy = s.get()
    return s, vx, vy


def direct_sum(x, y, strengths, alpha, dmin2, dmax2):
    """O(n^2) d3 pair sum (manyBody.js leaf branch with every cell opened):
    v_i = sum_{j != i} d_ij * s_j * alpha / l_ij, l < dmin^2 -> sqrt(dmin^2 l),
    l >= dmax^2 and coincident pairs skipped (error-test.js:44-58); also
    sum_j |term| per node (a bound for the rounding of any summation order)."""
    n = len(x)
    vx, vy, mag = np.zeros(n), np.zeros(n), np.zeros(n)
    for a in range(0, n, 512):
        b = min(n, a + 512)
        dx = x[None, :] - x[a:b, None]
        dy = y[None, :] - y[a:b, None]
        l = dx * dx + dy * dy
        l = np.where(l < dmin2, np.sqrt(dmin2 * l), l)
        keep = (l > 0) & (l < dmax2) & (np.arange(n)[None, :] != np.arange(a, b)[:, None])
        with np.errstate(divide="ignore", invalid="ignore"):
            w = np.where(keep, strengths[None, :] * alpha / l, 0.0)
        vx[a:b] = (dx * w).sum(1)
        vy[a:b] = (dy * w).sum(1)
        mag[a:b] = (np.abs(dx * w) + np.abs(dy * w)).sum(1)
    return vx, vy, mag


def check_theta_zero_is_direct_sum(core, n=600):
    """theta = 0: w^2 / theta^2 is +inf for every cell, so no cell is
    accepted and the traversal sums every leaf: d3's O(n^2) direct sum, in
    tree order. Mixed-sign strengths; distanceMin 5 (its sqrt branch) and a
    finite distanceMax (skipped leaves) in the second case."""
    x, y = _layout(n, 3)
    r = lcg(9)
    strengths = np.array([-40 + 60 * r() for _ in range(n)])
    alpha = 0.3
    for dmin2, dmax2 in ((1.0, INF), (25.0, 150.0 ** 2)):
        s, vx, vy = _many_body(core, x, y, strengths, 0.0, dmin2, dmax2, alpha)
        s.close()
        ex, ey, mag = direct_sum(x, y, strengths, alpha, dmin2, dmax2)
        bound = 1e-12 * mag + 1e-300
        assert np.all(np.abs(vx - ex) <= bound), np.max(np.abs(vx - ex) / bound)
        assert np.all(np.abs(vy - ey) <= bound), np.max(np.abs(vy - ey) / bound)


def check_theta_infinite_sees_only_root(core, n=500):
    """theta = inf: w^2 / theta^2 = 0 < l, so every node accepts the root at
    once and sees the whole charge S at its weighted centroid (for strengths
    of one sign the weights are |s_i|): v_i = (c - p_i) S alpha / l_i with
    d3's distanceMin rule; with a finite distanceMax, nodes at l >= dmax^2
    from the centroid get nothing."""
    x, y = _layout(n, 5)
    r = lcg(13)
    strengths = np.array([-10 - 40 * r() for _ in range(n)])
    alpha, dmin2 = 0.7, 4.0
    w = np.abs(strengths)
    cx, cy, S = (w * x).sum() / w.sum(), (w * y).sum() / w.sum(), strengths.sum()
    dx, dy = cx - x, cy - y
    l = dx * dx + dy * dy
    dmax2 = float(np.median(l))
    for cap in (INF, dmax2):
        s, vx, vy = _many_body(core, x, y, strengths, INF, dmin2, cap, alpha)
        s.close()
        ll = np.where(l < dmin2, np.sqrt(dmin2 * l), l)
        far = l >= cap
        np.testing.assert_allclose(vx[~far], (dx * S * alpha / ll)[~far], rtol=1e-9, atol=0)
        np.testing.assert_allclose(vy[~far], (dy * S * alpha / ll)[~far], rtol=1e-9, atol=0)
        assert np.all(vx[far] == 0) and np.all(vy[far] == 0)
        assert far.sum() == (0 if cap == INF else n // 2)


def check_cell_charges(core, n=3000):
    """Accumulate (visitAfter): every cell's value is the sum of its points'
    strengths and, for strengths of one sign, (x, y) is their |s|-weighted
    centroid; a leaf's (x, y) is its chain head's position."""
    from treeview import read_tree
    x, y = _layout(n, 21)
    x[:40] = x[40:80]  # coincident chains
    y[:40] = y[40:80]
    r = lcg(4)
    strengths = np.array([-5 - 30 * r() for _ in range(n)])
    s, _, _ = _many_body(core, x, y, strengths, 0.81, 1.0, INF, 0.5)
    extent, cells, perm, m = read_tree(core, s.h)
    s.close()
    assert m == n and len(cells) > n
    w = np.abs(strengths)
    for c in cells:
        pts = perm[c["first"]:c["first"] + c["count"]].astype(np.int64)
        assert len(pts) > 0
        want = strengths[pts].sum()
        assert abs(c["value"] - want) <= 1e-12 * np.abs(strengths[pts]).sum()
        cx = (w[pts] * x[pts]).sum() / w[pts].sum()
        cy = (w[pts] * y[pts]).sum() / w[pts].sum()
        assert abs(c["x"] - cx) <= 1e-9 * (1 + abs(cx)) and abs(c["y"] - cy) <= 1e-9 * (1 + abs(cy))


def _decayed_alpha():
    return 1 + (0 - 1) * (1 - math.pow(0.001, 1 / 300))  # simulation.js:64, first tick


def check_exact_tie(core):
    """test/parity/scenarios.js:176-185, derived by hand. Points (0,0), (4,0),
    (2,16); d3's cover gives the extent [0,32]^2. (0,0) and (4,0) share the
    cell [0,8]^2 (charge -60 at (2,0)); from (2,16) with theta 0.5,
    w^2/theta^2 = 64/0.25 = 256 = l, so the cell is NOT accepted (strict <)
    and the two leaves (w = 4: 64 < 260) are accepted one by one. Nodes 0
    and 1 see each other with an exactly zero y delta and jiggle it."""
    t = _nodes([{"x": 0.0, "y": 0.0}, {"x": 4.0, "y": 0.0}, {"x": 2.0, "y": 16.0}], core)
    d3.forceSimulation(t).force("f", d3.forceManyBody().theta(0.5)).tick()
    a = _decayed_alpha()
    # node 2: leaves (0,0) then (4,0), accepted: v += x * value * alpha / l
    vx2 = 0.0 + -2.0 * -30.0 * a / 260.0
    vx2 = vx2 + 2.0 * -30.0 * a / 260.0
    vy2 = 0.0 + -16.0 * -30.0 * a / 260.0
    vy2 = vy2 + -16.0 * -30.0 * a / 260.0
    # node 0: leaf (4,0) (w = 4, l = 16: 64 < 16 fails, so a leaf pair whose
    # y delta is exactly 0: jiggle y with the first draw, l += y^2), then
    # leaf (2,16) (w = 16, l = 260: 1024 < 260 fails, a leaf pair):
    # w = s * alpha / l; v += x * w. Node 1 mirrors it with the second draw.
    r = lcg()
    w2 = -30.0 * a / 260.0
    jy = (r() - 0.5) * 1e-6
    w1 = -30.0 * a / (16.0 + jy * jy)
    vx0 = 0.0 + 4.0 * w1
    vx0 = vx0 + 2.0 * w2
    vy0 = 0.0 + jy * w1
    vy0 = vy0 + 16.0 * w2
    jy = (r() - 0.5) * 1e-6
    w1 = -30.0 * a / (16.0 + jy * jy)
    vx1 = 0.0 + -4.0 * w1
    vx1 = vx1 + -2.0 * w2
    vy1 = 0.0 + jy * w1
    vy1 = vy1 + 16.0 * w2
    want = [(0.0, 0.0, vx0, vy0), (4.0, 0.0, vx1, vy1), (2.0, 16.0, vx2, vy2)]
    for i, (px, py, wx, wy) in enumerate(want):
        wx, wy = wx * 0.6, wy * 0.6  # integrate: v *= 1 - velocityDecay
        assert t.vx[i] == wx and t.vy[i] == wy, (i, t.vx[i], wx, t.vy[i], wy)
        assert t.x[i] == px + wx and t.y[i] == py + wy
    # Accepting the tied cell (<= instead of <) would give node 2
    # vy = -16 * -60 * a / 256 * 0.6 instead.
    assert t.vy[2] != -16.0 * -60.0 * a / 256.0 * 0.6


def check_coincident_jiggle_order(core):
    """Two nodes at the origin, derived by hand. The tree is one leaf whose
    chain head is node 1 (the later insert). Node 0 sees a leaf headed by
    another node, node 1 a leaf with a next: each jiggles x then y
    (manyBody.js: (random() - 0.5) * 1e-6, l += x^2 ...), l < distanceMin^2
    gives l = sqrt(l), then v += x * (s alpha / l) for the other node. Draws
    are the simulation's LCG in node order: r1, r2 for node 0, r3, r4 for 1."""
    t = _nodes([{"x": 0.0, "y": 0.0}, {"x": 0.0, "y": 0.0}], core)
    sim = d3.forceSimulation(t).force("f", d3.forceManyBody())
    sim.tick()
    r = lcg()
    a = _decayed_alpha()
    for i in range(2):
        jx = (r() - 0.5) * 1e-6
        l = 0.0 + jx * jx
        jy = (r() - 0.5) * 1e-6
        l = l + jy * jy
        l = math.sqrt(1.0 * l)
        w = -30.0 * a / l
        vx, vy = (0.0 + jx * w) * 0.6, (0.0 + jy * w) * 0.6
        assert t.vx[i] == vx and t.vy[i] == vy, (i, t.vx[i], vx, t.vy[i], vy)
        assert t.x[i] == 0.0 + vx and t.y[i] == 0.0 + vy
    # The stream advanced by exactly four draws.
    assert sim.random_source()() == r()


# ---------------------------------------------------------------- link


def check_link_projection_identities(core):
    """Gauss-Seidel link (link.js:34-57): with d = q_t - q_s for the shifted
    endpoints q = p + v and l = (|d| - distance) / |d| * alpha * k, the
    update q_t -= d l b, q_s += d l (1 - b) leaves an isolated edge at length
    |d| - alpha k (|d| - distance) and does not move the bias-weighted point
    (1 - b) q_t + b q_s. On a path whose edges run in order at alpha = k = 1
    the last edge ends exactly at its distance (it runs after every other)
    and the first does not (the second moved its shared node)."""
    n = 400
    m = n // 2
    x, y = _layout(n, 17)
    r = lcg(2)
    vx = np.array([r() - 0.5 for _ in range(n)])
    vy = np.array([r() - 0.5 for _ in range(n)])
    bias = np.array([0.1 + 0.8 * r() for _ in range(m)])
    dist = np.array([5 + 50 * r() for _ in range(m)])
    k = np.array([0.2 + 0.8 * r() for _ in range(m)])
    alpha = 0.6
    s = Raw(core, n)
    s.put(x, y, vx, vy)
    fh = s.force(C.LINK, m)
    s.u32(fh, C.SOURCE, m)[:] = np.arange(0, n, 2)
    s.u32(fh, C.TARGET, m)[:] = np.arange(1, n, 2)
    s.f64(fh, C.F0, m)[:] = k
    s.f64(fh, C.F1, m)[:] = dist
    s.f64(fh, C.F2, m)[:] = bias
    s.scalar(fh, C.ITERATIONS, 1)
    assert s.run(fh, alpha) == 0
    _, _, nvx, nvy = s.get()
    s.close()
    ox, oy = x + vx, y + vy
    before = np.hypot(ox[1::2] - ox[0::2], oy[1::2] - oy[0::2])
    qx, qy = x + nvx, y + nvy
    sx, sy, tx, ty = qx[0::2], qy[0::2], qx[1::2], qy[1::2]
    np.testing.assert_allclose(np.hypot(tx - sx, ty - sy), before - alpha * k * (before - dist), rtol=1e-12)
    for a, b in (((1 - bias) * tx + bias * sx, (1 - bias) * ox[1::2] + bias * ox[0::2]),
                 ((1 - bias) * ty + bias * sy, (1 - bias) * oy[1::2] + bias * oy[0::2])):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10)
    # A path 0-1-2-...: edges in order.
    p = 50
    s = Raw(core, p)
    s.put(x[:p], y[:p])
    fh = s.force(C.LINK, p - 1)
    s.u32(fh, C.SOURCE, p - 1)[:] = np.arange(p - 1)
    s.u32(fh, C.TARGET, p - 1)[:] = np.arange(1, p)
    s.f64(fh, C.F0, p - 1)[:] = 1.0
    s.f64(fh, C.F1, p - 1)[:] = 20.0
    s.f64(fh, C.F2, p - 1)[:] = 0.5
    s.scalar(fh, C.ITERATIONS, 1)
    assert s.run(fh, 1.0) == 0
    _, _, nvx, nvy = s.get()
    s.close()
    qx, qy = x[:p] + nvx, y[:p] + nvy
    assert abs(math.hypot(qx[p - 1] - qx[p - 2], qy[p - 1] - qy[p - 2]) - 20.0) < 1e-9
    assert abs(math.hypot(qx[1] - qx[0], qy[1] - qy[0]) - 20.0) > 1e-6


def check_center_identity(core):
    """forceCenter (center.js:9-12): every node moves by (mean - c) * strength,
    so the mean moves that fraction of the way to c and every pairwise
    offset is kept (up to rounding)."""
    n = 1000
    x, y = _layout(n, 8)
    x += 300.0
    s = Raw(core, n)
    s.put(x, y)
    fh = s.force(C.CENTER)
    s.scalar(fh, C.CX, 40.0)
    s.scalar(fh, C.CY, -25.0)
    s.scalar(fh, C.STRENGTH, 0.7)
    assert s.run(fh, 0.5) == 0
    nx, ny, nvx, nvy = s.get()
    s.close()
    for old, new, c in ((x, nx, 40.0), (y, ny, -25.0)):
        shift = (old.mean() - c) * 0.7
        np.testing.assert_allclose(new, old - shift, rtol=0, atol=1e-9)
    assert np.all(nvx == 0) and np.all(nvy == 0)


# ---------------------------------------------------------------- error-test.js


def error_scenarios():
    """test/fast/error-test.js:23-30 at scale 1, inputs from its LCG."""
    def rng(seed):
        return lcg(seed)

    def uniform(r):
        return [{"x": 1000 * (r() - 0.5), "y": 1000 * (r() - 0.5)} for _ in range(4000)]

    def clusters(r):
        out = []
        for i in range(6000):
            c = i % 3
            cx, cy = (0, 400, -300)[c], (0, 100, 350)[c]
            a = r() * 2 * math.pi
            d = 40 * math.sqrt(-2 * math.log(1 - r()))
            out.append({"x": cx + d * math.cos(a), "y": cy + d * math.sin(a)})
        return out

    return {
        "uniform 4000": lambda: uniform(rng(11)),
        "three clusters 6000": lambda: clusters(rng(11)),
        "phyllotaxis 8000": lambda: [{} for _ in range(8000)],
        "grid 2500 with exact zero deltas": lambda: [{"x": 12.0 * (i % 50), "y": 12.0 * (i // 50)} for i in range(2500)],
        "dense sub-unit cluster 3000": lambda: [{"x": 20 * r_(), "y": 20 * r_()} for r_ in [rng(11)] for _ in range(3000)],
    }


def one_tick_forces(core, dicts, fast):
    """error-test.js:33-38: one tick of manyBody alone with velocityDecay 0,
    so vx, vy are the force sums; returns (x0, y0, vx, vy, alpha)."""
    t = _nodes(dicts, core)
    sim = d3.forceSimulation(t).velocity_decay(0).force("charge", d3.forceManyBody())
    if fast:
        sim.precision("fast")
    x0, y0 = t.x.copy(), t.y.copy()
    sim.tick()
    return x0, y0, t.vx.copy(), t.vy.copy(), sim.alpha()


def rms_error(vx, vy, ref, zero_delta):  # error-test.js:60-69
    fx, fy = ref
    ex = np.zeros_like(vx) if zero_delta else vx - fx
    ey = vy - fy
    return math.sqrt(float((ex * ex + ey * ey).sum() / (fx * fx + fy * fy).sum()))


def error_gate(core, name, make):
    """error-test.js:71-81 for one scenario: (err_bh, err_fast)."""
    x0, y0, bvx, bvy, alpha = one_tick_forces(core, make(), False)
    d = make()
    for i, p in enumerate(d):
        p["x"], p["y"] = float(x0[i]), float(y0[i])
    fx0, fy0, fvx, fvy, alpha_f = one_tick_forces(core, d, True)
    assert alpha_f == alpha
    assert bitwise(fx0, x0) and bitwise(fy0, y0)
    # error-test.js:44-58: coincident pairs skipped, l < 1 -> sqrt(l)
    n = len(x0)
    ex, ey, _ = direct_sum(x0, y0, np.full(n, -30.0), alpha, 1.0, INF)
    grid = name.startswith("grid")
    return rms_error(bvx, bvy, (ex, ey), grid), rms_error(fvx, fvy, (ex, ey), grid)
