"""Node and link tables: the host's structure-of-arrays stand-in for d3's
arrays of node and link objects.

A d3 node is an object with x, y, vx, vy and optional fx, fy; ``fx == null``
means "not fixed" while a numeric fx (even NaN) fixes the node
(src/core.js:129). ``Nodes`` keeps those fields as float64 arrays plus
``fx_set``/``fy_set`` masks, so copy-in and copy-out are array copies.
"""

import numpy as np


class Nodes:
    """n nodes; unset fields are NaN (d3's ``undefined``)."""

    def __init__(self, n=0, x=None, y=None, vx=None, vy=None, fx=None, fy=None):
        n = int(n)
        nan = lambda: np.full(n, np.nan)  # noqa: E731
        self.x = nan() if x is None else np.array(x, dtype=np.float64)
        self.y = nan() if y is None else np.array(y, dtype=np.float64)
        self.vx = nan() if vx is None else np.array(vx, dtype=np.float64)
        self.vy = nan() if vy is None else np.array(vy, dtype=np.float64)
        self.fx = nan()
        self.fy = nan()
        self.fx_set = np.zeros(n, dtype=bool)
        self.fy_set = np.zeros(n, dtype=bool)
        if fx is not None:
            self.fix(fx=fx)
        if fy is not None:
            self.fix(fy=fy)
        self.core = None  # the core whose Context holds these nodes

    def __len__(self):
        return len(self.x)

    @property
    def index(self):
        return np.arange(len(self), dtype=np.int64)

    def fix(self, fx=None, fy=None, where=None):
        """Sets fx (and/or fy) for the nodes in ``where`` (all by default);
        a NaN entry in a masked array of values still counts as set."""
        sel = slice(None) if where is None else where
        if fx is not None:
            self.fx[sel] = fx
            self.fx_set[sel] = True
        if fy is not None:
            self.fy[sel] = fy
            self.fy_set[sel] = True

    @classmethod
    def from_dicts(cls, dicts):
        n = len(dicts)
        t = cls(n)
        for i, d in enumerate(dicts):
            for k in ("x", "y", "vx", "vy"):
                v = d.get(k)
                if v is not None:
                    getattr(t, k)[i] = v
            if d.get("fx") is not None:
                t.fx[i] = d["fx"]
                t.fx_set[i] = True
            if d.get("fy") is not None:
                t.fy[i] = d["fy"]
                t.fy_set[i] = True
        return t

    def to_dicts(self):
        out = []
        for i in range(len(self)):
            d = {"index": i, "x": float(self.x[i]), "y": float(self.y[i]),
                 "vx": float(self.vx[i]), "vy": float(self.vy[i])}
            if self.fx_set[i]:
                d["fx"] = float(self.fx[i])
            if self.fy_set[i]:
                d["fy"] = float(self.fy[i])
            out.append(d)
        return out

    def state(self):
        """x, y, vx, vy stacked as one (4, n) float64 array (a copy)."""
        return np.stack([self.x, self.y, self.vx, self.vy]).copy()


class Links:
    """m links as node indices (d3 resolves ids to node objects in
    link.js initialize; here they are already indices)."""

    def __init__(self, source, target):
        self.source = np.ascontiguousarray(source, dtype=np.int64)
        self.target = np.ascontiguousarray(target, dtype=np.int64)
        if self.source.shape != self.target.shape:
            raise ValueError("source and target differ in length")

    def __len__(self):
        return len(self.source)
