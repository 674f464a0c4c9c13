"""An independent pure-Python restatement of d3-force 3.0.0 and d3-quadtree
3.0.1 as published (the algorithms the reference claims bit parity with,
README.md:36-46), used to pin the oracle's exact mode. It follows d3's own
structure, not the reference core's: an insertion-built quadtree
(d3-quadtree add.js / cover.js / addAll), visit and visitAfter with explicit
quad stacks, forceManyBody's accumulate/apply, forceLink's Gauss-Seidel pass,
forceCenter, forceX/Y/Radial and the simulation tick with integrate. Nodes are
dicts, as d3's node objects. Small inputs only (pure Python).
"""

import math

INF = float("inf")


def jiggle(random):  # d3-force jiggle.js
    return (random() - 0.5) * 1e-6


class Leaf:
    __slots__ = ("data", "next", "x", "y", "value", "r", "kx", "ky")

    def __init__(self, data):
        self.data = data
        self.next = None


class Internal(list):
    """A d3 internal node: an array of 4 children (None when empty)."""

    def __init__(self):
        super().__init__([None, None, None, None])
        self.x = self.y = self.value = 0.0


class Quadtree:
    def __init__(self):
        self.x0 = self.y0 = self.x1 = self.y1 = math.nan
        self.root = None

    def cover(self, x, y):  # d3-quadtree cover.js
        if math.isnan(x) or math.isnan(y):
            return self
        x0, y0, x1, y1 = self.x0, self.y0, self.x1, self.y1
        if math.isnan(x0):
            x0 = float(math.floor(x)) if math.isfinite(x) else x
            x1 = x0 + 1
            y0 = float(math.floor(y)) if math.isfinite(y) else y
            y1 = y0 + 1
        else:
            z = x1 - x0
            if not z:  # `x1 - x0 || 1`
                z = 1.0
            guard = 0
            while (x0 > x or x >= x1 or y0 > y or y >= y1) and guard < 2200:
                i = (int(y < y0) << 1) | int(x < x0)
                z *= 2
                if i == 0:
                    x1, y1 = x0 + z, y0 + z
                elif i == 1:
                    x0, y1 = x1 - z, y0 + z
                elif i == 2:
                    x1, y0 = x0 + z, y1 - z
                else:
                    x0, y0 = x1 - z, y1 - z
                guard += 1
        self.x0, self.y0, self.x1, self.y1 = x0, y0, x1, y1
        return self

    def add(self, x, y, d):  # d3-quadtree add.js
        if math.isnan(x) or math.isnan(y):
            return
        leaf = Leaf(d)
        leaf.kx, leaf.ky = x, y  # the accessor's coordinates (add.js reads them through tree._x, tree._y)
        node = self.root
        x0, y0, x1, y1 = self.x0, self.y0, self.x1, self.y1
        if node is None:
            self.root = leaf
            return
        parent, i = None, 0
        while isinstance(node, Internal):
            xm = (x0 + x1) / 2
            right = x >= xm
            if right:
                x0 = xm
            else:
                x1 = xm
            ym = (y0 + y1) / 2
            bottom = y >= ym
            if bottom:
                y0 = ym
            else:
                y1 = ym
            parent = node
            i = (int(bottom) << 1) | int(right)
            node = node[i]
            if node is None:
                parent[i] = leaf
                return
        xp, yp = node.kx, node.ky
        if x == xp and y == yp:
            leaf.next = node
            if parent is not None:
                parent[i] = leaf
            else:
                self.root = leaf
            return
        while True:
            if parent is not None:
                parent[i] = Internal()
                parent = parent[i]
            else:
                self.root = Internal()
                parent = self.root
            xm = (x0 + x1) / 2
            right = x >= xm
            if right:
                x0 = xm
            else:
                x1 = xm
            ym = (y0 + y1) / 2
            bottom = y >= ym
            if bottom:
                y0 = ym
            else:
                y1 = ym
            i = (int(bottom) << 1) | int(right)
            j = (int(yp >= ym) << 1) | int(xp >= xm)
            if i != j:
                break
        parent[j] = node
        parent[i] = leaf

    def add_all(self, data, fx=None, fy=None):  # d3-quadtree addAll
        x0 = y0 = INF
        x1 = y1 = -INF
        pts = []
        for d in data:
            x = fx(d) if fx else d["x"]
            y = fy(d) if fy else d["y"]
            if math.isnan(x) or math.isnan(y):
                pts.append(None)
                continue
            pts.append((x, y))
            if x < x0:
                x0 = x
            if x > x1:
                x1 = x
            if y < y0:
                y0 = y
            if y > y1:
                y1 = y
        if x0 > x1 or y0 > y1:
            return self
        self.cover(x0, y0).cover(x1, y1)
        for d, p in zip(data, pts):
            if p is not None:
                self.add(p[0], p[1], d)
        return self

    def visit(self, callback):  # d3-quadtree visit.js
        quads = []
        if self.root is not None:
            quads.append((self.root, self.x0, self.y0, self.x1, self.y1))
        while quads:
            node, x0, y0, x1, y1 = quads.pop()
            if not callback(node, x0, y0, x1, y1) and isinstance(node, Internal):
                xm, ym = (x0 + x1) / 2, (y0 + y1) / 2
                if node[3] is not None:
                    quads.append((node[3], xm, ym, x1, y1))
                if node[2] is not None:
                    quads.append((node[2], x0, ym, xm, y1))
                if node[1] is not None:
                    quads.append((node[1], xm, y0, x1, ym))
                if node[0] is not None:
                    quads.append((node[0], x0, y0, xm, ym))

    def visit_after(self, callback):  # d3-quadtree visitAfter.js
        quads, nxt = [], []
        if self.root is not None:
            quads.append((self.root, self.x0, self.y0, self.x1, self.y1))
        while quads:
            q = quads.pop()
            node, x0, y0, x1, y1 = q
            if isinstance(node, Internal):
                xm, ym = (x0 + x1) / 2, (y0 + y1) / 2
                if node[0] is not None:
                    quads.append((node[0], x0, y0, xm, ym))
                if node[1] is not None:
                    quads.append((node[1], xm, y0, x1, ym))
                if node[2] is not None:
                    quads.append((node[2], x0, ym, xm, y1))
                if node[3] is not None:
                    quads.append((node[3], xm, ym, x1, y1))
            nxt.append(q)
        while nxt:
            callback(*nxt.pop())


def _truthy(v):
    return v == v and v != 0


class ManyBody:
    """d3-force manyBody.js."""

    def __init__(self, strength=-30.0, theta=0.9, distance_min=1.0, distance_max=INF):
        self.strength = strength
        self.theta2 = theta * theta
        self.distance_min2 = distance_min * distance_min
        self.distance_max2 = distance_max * distance_max

    def initialize(self, nodes, random):
        self.nodes = nodes
        self.random = random
        self.strengths = [0.0] * len(nodes)
        for i, node in enumerate(nodes):
            s = self.strength
            self.strengths[node["index"]] = float(s(node, i) if callable(s) else s)

    def __call__(self, alpha):
        tree = Quadtree().add_all(self.nodes)
        strengths = self.strengths

        def accumulate(quad, *_):
            strength = 0.0
            if isinstance(quad, Internal):
                x = y = weight = 0.0
                for q in quad:
                    if q is not None:
                        c = abs(q.value)
                        if _truthy(c):
                            strength += q.value
                            weight += c
                            x += c * q.x
                            y += c * q.y
                quad.x = _div(x, weight)
                quad.y = _div(y, weight)
            else:
                q = quad
                q.x = q.data["x"]
                q.y = q.data["y"]
                while q is not None:
                    strength += strengths[q.data["index"]]
                    q = q.next
            quad.value = strength

        tree.visit_after(accumulate)
        random = self.random
        theta2, dmin2, dmax2 = self.theta2, self.distance_min2, self.distance_max2
        for node in self.nodes:
            def apply(quad, x1, _y0, x2, _y1, node=node):
                if not _truthy(quad.value):
                    return True
                x = quad.x - node["x"]
                y = quad.y - node["y"]
                w = x2 - x1
                l = x * x + y * y
                if _div(w * w, theta2) < l:
                    if l < dmax2:
                        if x == 0:
                            x = jiggle(random)
                            l += x * x
                        if y == 0:
                            y = jiggle(random)
                            l += y * y
                        if l < dmin2:
                            l = math.sqrt(dmin2 * l)
                        node["vx"] += _div(x * quad.value * alpha, l)
                        node["vy"] += _div(y * quad.value * alpha, l)
                    return True
                if isinstance(quad, Internal) or l >= dmax2:
                    return None
                if quad.data is not node or quad.next is not None:
                    if x == 0:
                        x = jiggle(random)
                        l += x * x
                    if y == 0:
                        y = jiggle(random)
                        l += y * y
                    if l < dmin2:
                        l = math.sqrt(dmin2 * l)
                q = quad
                while q is not None:
                    if q.data is not node:
                        wj = _div(strengths[q.data["index"]] * alpha, l)
                        node["vx"] += x * wj
                        node["vy"] += y * wj
                    q = q.next
                return None

            tree.visit(apply)


def _div(a, b):
    """IEEE division as JS does it (Python raises on x / 0)."""
    try:
        return a / b
    except ZeroDivisionError:
        if a != a or a == 0:
            return math.nan
        neg = (math.copysign(1.0, a) < 0) != (math.copysign(1.0, b) < 0)
        return -INF if neg else INF


def _sqrt(v):
    if v != v or v < 0:
        return math.nan
    return math.sqrt(v)


class Link:
    """d3-force link.js with node indices for source and target."""

    def __init__(self, links, strength=None, distance=30.0, iterations=1):
        self.links = links  # list of (source, target) node indices
        self.strength_fn = strength
        self.distance_fn = distance
        self.iterations = iterations

    def initialize(self, nodes, random):
        self.nodes = nodes
        self.random = random
        n, m = len(nodes), len(self.links)
        count = [0] * n
        for s, t in self.links:
            count[s] += 1
            count[t] += 1
        self.bias = [count[s] / (count[s] + count[t]) for s, t in self.links]
        if self.strength_fn is None:
            self.strengths = [1 / min(count[s], count[t]) for s, t in self.links]
        else:
            f = self.strength_fn
            self.strengths = [float(f(i) if callable(f) else f) for i in range(m)]
        f = self.distance_fn
        self.distances = [float(f(i) if callable(f) else f) for i in range(m)]

    def __call__(self, alpha):
        nodes, random = self.nodes, self.random
        k = 0
        while k < self.iterations:
            for i, (s, t) in enumerate(self.links):
                source, target = nodes[s], nodes[t]
                x = target["x"] + target["vx"] - source["x"] - source["vx"]
                if not _truthy(x):
                    x = jiggle(random)
                y = target["y"] + target["vy"] - source["y"] - source["vy"]
                if not _truthy(y):
                    y = jiggle(random)
                l = _sqrt(x * x + y * y)
                l = _div(l - self.distances[i], l) * alpha * self.strengths[i]
                x *= l
                y *= l
                b = self.bias[i]
                target["vx"] -= x * b
                target["vy"] -= y * b
                b = 1 - b
                source["vx"] += x * b
                source["vy"] += y * b
            k += 1


class Collide:
    """d3-force collide.js: per iteration a quadtree on x + vx, y + vy with
    each quad's radius from visitAfter(prepare), then for every node a visit
    that pushes apart the overlapping later nodes (data.index > node.index)."""

    def __init__(self, radius=1.0, strength=1.0, iterations=1):
        self.radius = radius
        self.strength = strength
        self.iterations = iterations

    def initialize(self, nodes, random):
        self.nodes = nodes
        self.random = random
        self.radii = [0.0] * len(nodes)
        for i, node in enumerate(nodes):
            r = self.radius
            self.radii[node["index"]] = float(r(node, i) if callable(r) else r)

    def __call__(self, alpha):
        radii, strength = self.radii, self.strength

        def prepare(quad, *_):
            if isinstance(quad, Leaf):
                quad.r = radii[quad.data["index"]]
                return
            quad.r = 0.0
            for q in quad:
                if q is not None and q.r > quad.r:
                    quad.r = q.r

        for _ in range(self.iterations):
            tree = Quadtree().add_all(self.nodes, lambda d: d["x"] + d["vx"], lambda d: d["y"] + d["vy"])
            tree.visit_after(prepare)
            for node in self.nodes:
                ri = radii[node["index"]]
                ri2 = ri * ri
                xi = node["x"] + node["vx"]
                yi = node["y"] + node["vy"]

                def apply(quad, x0, y0, x1, y1):
                    rj = quad.r
                    r = ri + rj
                    if isinstance(quad, Leaf):
                        data = quad.data
                        if data["index"] > node["index"]:
                            x = xi - data["x"] - data["vx"]
                            y = yi - data["y"] - data["vy"]
                            l = x * x + y * y
                            if l < r * r:
                                if x == 0:
                                    x = jiggle(self.random)
                                    l += x * x
                                if y == 0:
                                    y = jiggle(self.random)
                                    l += y * y
                                l = _sqrt(l)
                                l = _div(r - l, l) * strength
                                x *= l
                                y *= l
                                rj *= rj
                                r = _div(rj, ri2 + rj)
                                node["vx"] += x * r
                                node["vy"] += y * r
                                r = 1 - r
                                data["vx"] -= x * r
                                data["vy"] -= y * r
                        return False
                    return x0 > xi + r or x1 < xi - r or y0 > yi + r or y1 < yi - r

                tree.visit(apply)


class Center:
    """d3-force center.js."""

    def __init__(self, x=0.0, y=0.0, strength=1.0):
        self.x, self.y, self.strength = x, y, strength

    def initialize(self, nodes, random):
        self.nodes = nodes

    def __call__(self, alpha):
        n = len(self.nodes)
        sx = sy = 0.0
        for node in self.nodes:
            sx += node["x"]
            sy += node["y"]
        sx = (_div(sx, n) - self.x) * self.strength
        sy = (_div(sy, n) - self.y) * self.strength
        for node in self.nodes:
            node["x"] -= sx
            node["y"] -= sy


class Axis:
    """d3-force x.js / y.js."""

    def __init__(self, key, target=0.0, strength=0.1):
        self.key, self.target, self.strength = key, target, strength

    def initialize(self, nodes, random):
        self.nodes = nodes
        self.z, self.s = [], []
        for i, node in enumerate(nodes):
            t = self.target
            z = float(t(node, i) if callable(t) else t)
            st = self.strength
            self.z.append(z)
            self.s.append(0.0 if math.isnan(z) else float(st(node, i) if callable(st) else st))

    def __call__(self, alpha):
        p, v = self.key, "v" + self.key
        for i, node in enumerate(self.nodes):
            node[v] += (self.z[i] - node[p]) * self.s[i] * alpha


class Radial:
    """d3-force radial.js."""

    def __init__(self, radius, x=0.0, y=0.0, strength=0.1):
        self.radius, self.x, self.y, self.strength = radius, x, y, strength

    def initialize(self, nodes, random):
        self.nodes = nodes
        self.r, self.s = [], []
        for i, node in enumerate(nodes):
            rr = self.radius
            r = float(rr(node, i) if callable(rr) else rr)
            st = self.strength
            self.r.append(r)
            self.s.append(0.0 if math.isnan(r) else float(st(node, i) if callable(st) else st))

    def __call__(self, alpha):
        for i, node in enumerate(self.nodes):
            dx = node["x"] - self.x
            if not _truthy(dx):
                dx = 1e-6
            dy = node["y"] - self.y
            if not _truthy(dy):
                dy = 1e-6
            r = _sqrt(dx * dx + dy * dy)
            k = _div((self.r[i] - r) * self.s[i] * alpha, r)
            node["vx"] += dx * k
            node["vy"] += dy * k


class Simulation:
    """d3-force simulation.js: initializeNodes and tick."""

    def __init__(self, nodes, random):
        self.nodes = nodes
        self.random = random
        self.alpha = 1.0
        self.alpha_decay = 1 - math.pow(0.001, 1 / 300)
        self.alpha_target = 0.0
        self.velocity_decay = 0.6
        self.forces = []
        for i, node in enumerate(nodes):
            node["index"] = i
            if node.get("fx") is not None:
                node["x"] = node["fx"]
            if node.get("fy") is not None:
                node["y"] = node["fy"]
            x, y = node.get("x", math.nan), node.get("y", math.nan)
            if x is None or y is None or math.isnan(x) or math.isnan(y):
                radius = 10 * math.sqrt(0.5 + i)
                angle = i * (math.pi * (3 - math.sqrt(5)))
                node["x"] = radius * math.cos(angle)
                node["y"] = radius * math.sin(angle)
            vx, vy = node.get("vx", math.nan), node.get("vy", math.nan)
            if vx is None or vy is None or math.isnan(vx) or math.isnan(vy):
                node["vx"] = node["vy"] = 0.0

    def force(self, f):
        f.initialize(self.nodes, self.random)
        self.forces.append(f)
        return self

    def tick(self, iterations=1):
        for _ in range(iterations):
            self.alpha += (self.alpha_target - self.alpha) * self.alpha_decay
            for f in self.forces:
                f(self.alpha)
            for node in self.nodes:
                if node.get("fx") is None:
                    node["vx"] *= self.velocity_decay
                    node["x"] += node["vx"]
                else:
                    node["x"] = node["fx"]
                    node["vx"] = 0.0
                if node.get("fy") is None:
                    node["vy"] *= self.velocity_decay
                    node["y"] += node["vy"]
                else:
                    node["y"] = node["fy"]
                    node["vy"] = 0.0
        return self
