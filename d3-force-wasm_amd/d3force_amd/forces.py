"""Native force facades, mirroring the reference's src/manyBody.js,
src/link.js, src/center.js, src/collide.js, src/x.js, src/y.js and
src/radial.js: accessors are evaluated on ``initialize`` into the core's
parameter arrays, scalars are pushed before every run, and ``force(alpha)``
runs the kernel through its Slot.

Accessors may be a number, a numpy array (one value per node or link), or a
callable ``f(i) -> number`` / ``f(index_array) -> array``.
"""

import numpy as np

from .core import (Slot, MANY_BODY, LINK, CENTER, COLLIDE, RADIAL, X, Y,
                   THETA2, DISTANCE_MIN2, DISTANCE_MAX2, ITERATIONS, STRENGTH,
                   CX, CY, F0, F1, F2, SOURCE, TARGET)


def _values(accessor, n):
    """Evaluates a d3 accessor for n items as float64 (``+accessor(d, i)``)."""
    if callable(accessor):
        idx = np.arange(n)
        try:
            v = accessor(idx)
            v = np.asarray(v, dtype=np.float64)
            if v.shape == (n,):
                return v
        except Exception:  # an accessor of one item at a time
            pass
        return np.array([float(accessor(i)) for i in range(n)], dtype=np.float64)
    a = np.asarray(accessor, dtype=np.float64)
    if a.ndim == 0:
        return np.full(n, float(a))
    if a.shape != (n,):
        raise ValueError("accessor array has %s entries, expected %d" % (a.shape, n))
    return a.copy()


class _Force:
    """Callable force object: ``force(alpha)`` plus ``initialize``."""

    def __init__(self, kind):
        self.slot = Slot(kind)
        self.nodes = None
        self.random = None

    def initialize(self, nodes, random=None):
        self.nodes = nodes
        self.random = random
        self._initialize()

    def _initialize(self):
        pass

    def prepare(self):
        """The scalars a run sets before the core runs the force."""


class ManyBody(_Force):
    """manyBody.js:4-54."""

    def __init__(self):
        super().__init__(MANY_BODY)
        self._strength = -30.0
        self._distance_min2 = 1.0
        self._distance_max2 = float("inf")
        self._theta2 = 0.81

    def __call__(self, alpha):  # manyBody.js:14-17
        self.push_scalars()
        self.slot.run(self.nodes, alpha, self.random)

    def prepare(self):
        self.push_scalars()

    def push_scalars(self):
        """theta^2, distanceMin^2 and distanceMax^2 into the core's force."""
        s = self.slot
        s.scalar(THETA2, self._theta2)
        s.scalar(DISTANCE_MIN2, self._distance_min2)
        s.scalar(DISTANCE_MAX2, self._distance_max2)

    def _initialize(self):  # manyBody.js:20-27
        if self.nodes is None:
            return
        n = len(self.nodes)
        strengths = _values(self._strength, n)
        s = self.slot
        s.bind(self.nodes)
        s.count(n)
        s.f64(F0, n)[:] = strengths

    def strength(self, v):
        self._strength = v
        self._initialize()
        return self

    def distance_min(self, v):
        self._distance_min2 = float(v) * float(v)
        return self

    def distance_max(self, v):
        self._distance_max2 = float(v) * float(v)
        return self

    def theta(self, v):
        self._theta2 = float(v) * float(v)
        return self


class Link(_Force):
    """link.js:13-121 over a ``Links`` table of node indices."""

    def __init__(self, links):
        super().__init__(LINK)
        self.links = links
        self._strength = None  # None: d3's default 1 / min(count)
        self._distance = 30.0
        self._iterations = 1
        self.size = 0

    def __call__(self, alpha):  # link.js:34-37
        self.prepare()
        self.slot.run(self.nodes, alpha, self.random)

    def prepare(self):
        self.slot.scalar(ITERATIONS, self._iterations)

    def _initialize(self):  # link.js:39-67
        if self.nodes is None:
            return
        n = len(self.nodes)
        src = self.links.source
        tgt = self.links.target
        m = len(src)
        if m and (src.min() < 0 or tgt.min() < 0 or src.max() >= n or tgt.max() >= n):
            raise ValueError("node not found")
        count = (np.bincount(src, minlength=n) + np.bincount(tgt, minlength=n)).astype(np.float64)
        self.count = count
        bias = count[src] / (count[src] + count[tgt])
        s = self.slot
        s.bind(self.nodes)
        s.count(m)
        self.size = m
        s.u32(SOURCE, m)[:] = src
        s.u32(TARGET, m)[:] = tgt
        s.f64(F2, m)[:] = bias
        self._initialize_strength()
        self._initialize_distance()

    def _initialize_strength(self):  # link.js:69-76, default :30-32
        if self.nodes is None:
            return
        src, tgt = self.links.source, self.links.target
        if self._strength is None:
            v = 1.0 / np.minimum(self.count[src], self.count[tgt])
        else:
            v = _values(self._strength, len(src))
        self.slot.f64(F0, self.size)[:] = v[: self.size]

    def _initialize_distance(self):  # link.js:78-85
        if self.nodes is None:
            return
        v = _values(self._distance, len(self.links.source))
        self.slot.f64(F1, self.size)[:] = v[: self.size]

    def set_links(self, links):  # link.js:98-100, d3's link.links(_)
        if not hasattr(links, "source"):
            from .nodes import Links
            links = Links(*links)
        self.links = links
        self._initialize()
        return self

    def iterations(self, k):
        self._iterations = float(k)
        return self

    def strength(self, v):
        self._strength = v
        self._initialize_strength()
        return self

    def distance(self, v):
        self._distance = v
        self._initialize_distance()
        return self


class Center(_Force):
    """center.js:3-34."""

    def __init__(self, x=0.0, y=0.0):
        super().__init__(CENTER)
        self._x = 0.0 if x is None else float(x)
        self._y = 0.0 if y is None else float(y)
        self._strength = 1.0

    def __call__(self, alpha=None):  # center.js:9-12
        self.prepare()
        self.slot.run(self.nodes, 0.0)

    def prepare(self):
        s = self.slot
        s.scalar(CX, self._x)
        s.scalar(CY, self._y)
        s.scalar(STRENGTH, self._strength)

    def initialize(self, nodes, random=None):  # center.js:16-19
        self.nodes = nodes
        self.slot.bind(nodes)

    def strength(self, v):
        self._strength = float(v)
        return self


class Collide(_Force):
    """collide.js:4-50."""

    def __init__(self, radius=1.0):
        super().__init__(COLLIDE)
        self._radius = 1.0 if radius is None else radius
        self._strength = 1.0
        self._iterations = 1

    def __call__(self, alpha=None):  # collide.js:14-17
        s = self.slot
        s.scalar(STRENGTH, self._strength)
        s.scalar(ITERATIONS, self._iterations)
        s.run(self.nodes, 0.0, self.random)

    def _initialize(self):  # collide.js:20-27
        if self.nodes is None:
            return
        n = len(self.nodes)
        radii = _values(self._radius, n)
        s = self.slot
        s.bind(self.nodes)
        s.count(n)
        s.f64(F0, n)[:] = radii

    def radius(self, v):
        self._radius = v
        self._initialize()
        return self

    def strength(self, v):
        self._strength = float(v)
        return self

    def iterations(self, k):
        self._iterations = float(k)
        return self


class Axis(_Force):
    """x.js / y.js: ``v += (target - p) * strength * alpha``."""

    def __init__(self, kind, target=0.0):
        super().__init__(kind)
        self._target = 0.0 if target is None else target
        self._strength = 0.1

    def __call__(self, alpha):
        self.slot.run(self.nodes, alpha)

    def initialize(self, nodes, random=None):
        self.nodes = nodes
        self._initialize()

    def _initialize(self):  # x.js:14-24
        if self.nodes is None:
            return
        n = len(self.nodes)
        tz = _values(self._target, n)
        st = np.where(np.isnan(tz), 0.0, _values(self._strength, n))
        s = self.slot
        s.bind(self.nodes)
        s.count(n)
        s.f64(F0, n)[:] = st
        s.f64(F1, n)[:] = tz

    def strength(self, v):
        self._strength = v
        self._initialize()
        return self


class Radial(_Force):
    """radial.js:4-56."""

    def __init__(self, radius, x=0.0, y=0.0):
        super().__init__(RADIAL)
        self._radius = radius
        self._x = 0.0 if x is None else float(x)
        self._y = 0.0 if y is None else float(y)
        self._strength = 0.1

    def __call__(self, alpha):
        self.prepare()
        self.slot.run(self.nodes, alpha)

    def prepare(self):
        s = self.slot
        s.scalar(CX, self._x)
        s.scalar(CY, self._y)

    def initialize(self, nodes, random=None):
        self.nodes = nodes
        self._initialize()

    def _initialize(self):  # radial.js:20-31
        if self.nodes is None:
            return
        n = len(self.nodes)
        radii = _values(self._radius, n)
        st = np.where(np.isnan(radii), 0.0, _values(self._strength, n))
        s = self.slot
        s.bind(self.nodes)
        s.count(n)
        s.f64(F0, n)[:] = st
        s.f64(F1, n)[:] = radii


def forceManyBody():
    return ManyBody()


def forceLink(links=None):
    if links is None:
        links = ([], [])
    if not hasattr(links, "source"):
        from .nodes import Links
        src, tgt = links
        links = Links(src, tgt)
    return Link(links)


def forceCenter(x=None, y=None):
    return Center(x, y)


def forceCollide(radius=None):
    return Collide(radius)


def forceX(x=None):
    return Axis(X, x)


def forceY(y=None):
    return Axis(Y, y)


def forceRadial(radius, x=None, y=None):
    return Radial(radius, x, y)
