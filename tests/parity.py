"""Scenario specs of the reference's parity corpus (test/parity/scenarios.js)
and builders that run one spec on d3ref (the independent restatement of
d3-force) or on any core through the host facade."""

import math

import numpy as np
import d3force_amd as d3
import d3ref
from scenarios import generator

INF = float("inf")


def xorshift(seed):  # test/parity/scenarios.js:18-24
    s = [(seed | 0) or 1]

    def r():
        v = s[0] & 0xFFFFFFFF
        v ^= (v << 13) & 0xFFFFFFFF
        v ^= v >> 17
        v ^= (v << 5) & 0xFFFFFFFF
        s[0] = v
        return v / 4294967296

    return r


def random_nodes(n, r, spread=500):
    return [{"x": (r() - 0.5) * spread, "y": (r() - 0.5) * spread} for _ in range(n)]


def random_links(n, m, r):
    return [(math.floor(r() * n), math.floor(r() * n)) for _ in range(m)]


def defaults(nodes, links):
    return nodes, [("charge", {}), ("link", {"links": links}), ("center", {})]


def scale_free_links(n, r):  # test/parity/scenarios.js:38-47
    links, ends = [], [0]
    for i in range(1, n):
        k = 2 if r() < 0.3 else 1
        while k > 0:
            j = ends[math.floor(r() * len(ends))]
            links.append((i, j))
            ends.extend([i, j])
            k -= 1
    return links


def with_collide(spec, radius):
    nodes, forces = spec
    return nodes, forces + [("collide", {"radius": radius})]


def _scenarios():
    S = {}
    for n in (10, 100, 300):
        S["default-forces-%d" % n] = lambda r, n=n: defaults([{} for _ in range(n)], random_links(n, 2 * n, r))
    def scale_free(r):
        nodes = random_nodes(300, r)
        return with_collide(defaults(nodes, scale_free_links(300, r)), 4.0)
    S["scale-free-collide"] = scale_free
    S["grid"] = lambda r: defaults([{"x": float(i % 20), "y": float(i // 20)} for i in range(400)],
                                   [l for i in range(400) for l in
                                    ([(i, i + 1)] if i % 20 < 19 else []) + ([(i, i + 20)] if i < 380 else [])])
    S["all-at-origin"] = lambda r: with_collide(
        defaults([{"x": 0.0, "y": 0.0} for _ in range(60)], random_links(60, 60, r)), 2.0)
    S["coincident-groups"] = lambda r: with_collide(
        defaults([{"x": (i % 6) * 10.0, "y": (i % 4) * 10.0} for i in range(120)], random_links(120, 100, r)), 3.0)
    S["near-1e300"] = lambda r: defaults([{"x": 1e300 * (r() - 0.5), "y": 1e300 * (r() - 0.5)} for _ in range(50)],
                                         random_links(50, 40, r))
    S["near-duplicates"] = lambda r: with_collide(defaults(
        [{"x": i * 300.0, "y": i * 300.0} if i < 4 else
         {"x": 0.5 + math.floor(r() * 6) * 1e-12, "y": 0.25 + math.floor(r() * 6) * 3e-13} for i in range(80)],
        random_links(80, 40, r)), 1.0)

    def single(force, n=200):
        def make(r):
            params = force[1](r) if callable(force[1]) else force[1]
            return random_nodes(n, r), [(force[0], params)]
        return make

    S["only-center"] = single(("center", lambda r: {"x": 30 * r(), "y": -20 * r(), "strength": 0.7}))
    S["only-collide"] = single(("collide", {"radius": 12.0}))
    S["only-link"] = lambda r: (random_nodes(150, r), [("link", {"links": random_links(150, 300, r)})])
    S["only-many-body"] = single(("charge", {}))
    S["only-radial"] = single(("radial", lambda r: {"radius": 100.0, "x": 10 * r(), "y": 5.0}))
    S["only-x"] = single(("x", lambda r: {"target": 50 * r()}))
    S["only-y"] = single(("y", lambda r: {"target": -50 * r()}))
    S["nan-accessors"] = lambda r: (random_nodes(100, r), [
        ("charge", {"strength": lambda i: math.nan if i % 13 == 0 else -20.0}),
        ("collide", {"radius": lambda i: math.nan if i % 11 == 0 else 3.0}),
        ("x", {"target": lambda i: math.nan if i % 7 == 0 else 0.0}),
        ("radial", {"radius": lambda i: math.nan if i % 5 == 0 else 80.0})])
    S["nan-link-distance"] = lambda r: (random_nodes(60, r), [
        ("link", {"links": random_links(60, 50, r), "distance": lambda i: math.nan if i == 3 else 20.0})])
    S["many-body-exact-tie"] = lambda r: ([{"x": 0.0, "y": 0.0}, {"x": 4.0, "y": 0.0}, {"x": 2.0, "y": 16.0}],
                                          [("charge", {"theta": 0.5})])

    def inf_links(r):
        nodes = [({"x": -INF if i % 2 == 0 else INF, "y": INF} if i < 6 else d)
                 for i, d in enumerate(random_nodes(30, r))]
        links = random_links(30, 30, r) + [(0, 2), (1, 3), (4, 5)]
        return nodes, [("link", {"links": links})]
    S["link-infinite-positions"] = inf_links
    S["many-body-theta-0"] = single(("charge", {"theta": 0.0}), 120)
    S["many-body-theta-infinity"] = single(("charge", {"theta": INF}), 120)
    S["many-body-distance-max"] = single(("charge", {"distance_max": 60.0, "distance_min": 5.0}), 300)

    def zero_links(r):
        nodes = [({"x": 0.0, "y": 0.0} if i < 10 else d) for i, d in enumerate(random_nodes(40, r))]
        links = random_links(40, 40, r) + [(i, (i + 1) % 10) for i in range(10)] + [(20, 20)]
        return nodes, [("link", {"links": links})]
    S["zero-length-links"] = zero_links
    S["iterations-and-parameters"] = lambda r: (random_nodes(200, r), [
        ("charge", {"strength": -15.0, "theta": 0.5}),
        ("link", {"links": random_links(200, 250, r), "iterations": 2, "distance": 15.0}),
        ("collide", {"radius": 6.0, "iterations": 3, "strength": 0.7}),
        ("sim", {"alpha_decay": 0.01, "alpha_target": 0.2, "velocity_decay": 0.25})])

    def nan_initial(r):
        nodes = []
        for i, d in enumerate(random_nodes(100, r)):
            nodes.append({"y": d["y"]} if i % 4 == 0 else ({"x": d["x"], "vx": math.nan} if i % 4 == 1 else d))
        return with_collide(defaults(nodes, random_links(100, 120, r)), 2.0)
    S["nan-initial-positions"] = nan_initial

    def xs(r):
        nodes, forces = with_collide(defaults([{"x": 0.0, "y": 0.0} for _ in range(40)], random_links(40, 30, r)), 2.0)
        return nodes, forces + [("random", {"seed": math.floor(r() * 2 ** 31)})]
    S["random-source-xorshift"] = xs
    S["pinned-nodes"] = lambda r: defaults(
        [({"x": 0.0, "y": 0.0, "fx": 1000 * (r() - 0.5), "fy": 1000 * (r() - 0.5)} if i % 10 == 0 else
          {"x": (r() - 0.5) * 500, "y": (r() - 0.5) * 500}) for i in range(200)],
        random_links(200, 400, r))
    return S


SCENARIOS = _scenarios()


def build_d3ref(nodes, forces):
    nodes = [dict(d) for d in nodes]
    rand = d3.lcg()
    for name, p in forces:
        if name == "random":
            rand = xorshift(p["seed"])
    sim = d3ref.Simulation(nodes, rand)
    for name, p in forces:
        if name == "sim":
            sim.alpha_decay = p["alpha_decay"]
            sim.alpha_target = p["alpha_target"]
            sim.velocity_decay = 1 - p["velocity_decay"]
    for name, p in forces:
        acc = lambda v: (lambda node, i: v(i)) if callable(v) else v  # noqa: E731
        if name == "charge":
            sim.force(d3ref.ManyBody(strength=acc(p.get("strength", -30.0)), theta=p.get("theta", 0.9),
                                     distance_min=p.get("distance_min", 1.0),
                                     distance_max=p.get("distance_max", INF)))
        elif name == "link":
            sim.force(d3ref.Link(p["links"], distance=p.get("distance", 30.0), iterations=p.get("iterations", 1)))
        elif name == "center":
            sim.force(d3ref.Center(p.get("x", 0.0), p.get("y", 0.0), p.get("strength", 1.0)))
        elif name in ("x", "y"):
            sim.force(d3ref.Axis(name, acc(p.get("target", 0.0))))
        elif name == "radial":
            sim.force(d3ref.Radial(acc(p["radius"]), p.get("x", 0.0), p.get("y", 0.0)))
        elif name == "collide":
            sim.force(d3ref.Collide(acc(p["radius"]), p.get("strength", 1.0), p.get("iterations", 1)))
    return sim


def build_facade(core, nodes, forces):
    table = d3.Nodes.from_dicts(nodes)
    sim = d3.forceSimulation(table, core=core)
    for name, p in forces:
        if name == "random":
            sim.random_source(xorshift(p["seed"]))
        if name == "sim":
            sim.alpha_decay(p["alpha_decay"]).alpha_target(p["alpha_target"]).velocity_decay(p["velocity_decay"])
    for k, (name, p) in enumerate(forces):
        if name == "charge":
            f = d3.forceManyBody().strength(p.get("strength", -30.0)).theta(p.get("theta", 0.9))
            f.distance_min(p.get("distance_min", 1.0)).distance_max(p.get("distance_max", INF))
        elif name == "link":
            src = [s for s, _ in p["links"]]
            tgt = [t for _, t in p["links"]]
            f = d3.forceLink((src, tgt)).distance(p.get("distance", 30.0)).iterations(p.get("iterations", 1))
        elif name == "center":
            f = d3.forceCenter(p.get("x", 0.0), p.get("y", 0.0)).strength(p.get("strength", 1.0))
        elif name == "x":
            f = d3.forceX(p.get("target", 0.0))
        elif name == "y":
            f = d3.forceY(p.get("target", 0.0))
        elif name == "radial":
            f = d3.forceRadial(p["radius"], p.get("x", 0.0), p.get("y", 0.0))
        elif name == "collide":
            rad = p["radius"]
            f = d3.forceCollide([rad(i) for i in range(len(nodes))] if callable(rad) else rad)
            f.strength(p.get("strength", 1.0)).iterations(p.get("iterations", 1))
        else:
            continue
        sim.force("%s%d" % (name, k), f)
    return sim, table


def same(a, b):
    """Object.is over arrays: -0 differs from 0, NaN equals NaN."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    eq = (a.view(np.uint64) == b.view(np.uint64)) | both_nan
    return bool(np.all(eq))


def d3ref_state(sim):
    return np.array([[d[k] for d in sim.nodes] for k in ("x", "y", "vx", "vy")], dtype=np.float64)


