"""Restatement of the v1 per-object reference solver (the timed "direbm/reference" CPU baseline).

Follows direbm/reference/simulator.py:30-198 and direbm/reference/grid.py:17-92 (2D, no obstacle):
collision, dispersion, greedy density-threshold control points, refinement with the soft_outer
spawn, and resampling with create-on-empty and gap-fill.  The algorithm is order-dependent
(control points are created greedily in the iteration order of a dict of cells), so every
container operation keeps the reference's insertion order: cells of a dict keyed by
floor(x / cell), items appended, emptied cells deleted, neighbour cells visited in
itertools.product order.  Pinned by golden vectors written from the reference package
(tests/golden/make_v1_golden.py).
"""

from __future__ import annotations

import math
from itertools import product

import numpy as np

from .constants import ALPHA, C2, CS2_2D, DX, KAPPA_HARD, KAPPA_SOFT, W2
from .physics import collide, equilibrium

EPS = 1e-8  # direbm/reference/simulator.py:24
SOFT_OFFSET = 2.0 * (1.0 - np.sqrt(3.0) / 2.0)  # simulator.py:27


class Item:
    """A moment / component / control point: position x (float64 (2,)) and fields."""

    __slots__ = ("x", "f", "i", "near", "kappa", "kind")

    def __init__(self, x, f=None, i=-1):
        self.x = x
        self.f = f
        self.i = i
        self.near = []
        self.kappa = 0
        self.kind = ""


class CellHash:
    """Dict-of-cells spatial hash with radius queries (grid.py:17-92)."""

    def __init__(self, cell):
        self.cell = float(cell)
        self.cells: dict = {}

    def key(self, x):
        return tuple(int(math.floor(float(v) / self.cell)) for v in x)

    def _keys_around(self, x, radius):
        r = int(math.ceil(radius / self.cell))
        base = self.key(x)
        for off in product(range(-r, r + 1), repeat=len(base)):
            yield tuple(b + o for b, o in zip(base, off))

    def insert(self, item):
        self.cells.setdefault(self.key(item.x), []).append(item)
        return item

    def near(self, x, radius):
        x = np.asarray(x, dtype=np.float64)
        r2 = radius * radius
        out = []
        for k in self._keys_around(x, radius):
            bucket = self.cells.get(k)
            if bucket:
                for it in bucket:
                    d = it.x - x
                    if float(d @ d) <= r2:
                        out.append(it)
        return out

    def remove_near(self, x, radius):
        x = np.asarray(x, dtype=np.float64)
        r2 = radius * radius
        for k in self._keys_around(x, radius):
            bucket = self.cells.get(k)
            if not bucket:
                continue
            keep = [it for it in bucket if float((it.x - x) @ (it.x - x)) > r2]
            if keep:
                self.cells[k] = keep
            else:
                del self.cells[k]

    def items(self):
        out = []
        for bucket in self.cells.values():
            out.extend(bucket)
        return out


class Simulator:
    """v1 DiReBM (2D D2Q7): Simulator(pos (M,2), f (M,7), tau, ...); .step(); .moments()."""

    def __init__(self, pos, f, tau, *, dx=DX, alpha=ALPHA, kappa_hard=KAPPA_HARD, kappa_soft=KAPPA_SOFT,
                 rho_rest=1.0, soft_mode="spawn"):
        pos = np.asarray(pos, np.float64).reshape(-1, 2)
        f = np.asarray(f, np.float64).reshape(-1, 7)
        self.moments_ = [Item(pos[k].copy(), f[k].copy()) for k in range(len(pos))]
        self.tau = float(tau)
        self.dx = float(dx)
        self.kappa_hard = int(kappa_hard)
        self.kappa_soft = int(kappa_soft)
        self.soft_mode = soft_mode
        self.r_thresh = self.dx / float(alpha) - EPS  # simulator.py:61
        self.f_rest = equilibrium(np.float64(rho_rest), np.zeros(2))
        self.comps = CellHash(self.dx)
        self.cps = CellHash(self.dx)
        self.iteration = 0

    def moments(self):
        if not self.moments_:
            return np.zeros((0, 2)), np.zeros((0, 7))
        return np.array([m.x for m in self.moments_]), np.stack([m.f for m in self.moments_])

    def step(self):
        """simulator.py:73-81 (no obstacle)."""
        if self.moments_:
            fs = collide(np.stack([m.f for m in self.moments_]), self.tau, C2, W2, CS2_2D)
            for m, fi in zip(self.moments_, fs):
                m.f = fi
        self._disperse()
        self._create()
        self._refine()
        self._resample()
        self.iteration += 1

    def _disperse(self):
        """simulator.py:92-103: each moment -> 7 components shifted dx along c_i."""
        self.comps.cells.clear()
        for m in self.moments_:
            for i in range(7):
                self.comps.insert(Item(m.x + C2[i] * self.dx, float(m.f[i]), i))
        self.moments_ = []

    def _create(self):
        """simulator.py:120-133: greedy control points at least dx/alpha apart, classified by kappa."""
        self.cps.cells.clear()
        for nu in self.comps.items():
            if self.cps.near(nu.x, self.r_thresh):
                continue
            p = self.cps.insert(Item(nu.x.copy(), np.zeros(7)))
            p.near = self.comps.near(p.x, self.dx + EPS)
            p.kappa = len({n.i for n in p.near})
            p.kind = "hard" if p.kappa <= self.kappa_hard else ("soft" if p.kappa <= self.kappa_soft else "inner")

    def _soft_point(self, p):
        """simulator.py:153-166."""
        if self.soft_mode == "off":
            return None
        c = np.zeros(2)
        for n in p.near:
            c += C2[n.i]
        nrm = np.linalg.norm(c)
        if nrm <= EPS:
            return None
        return p.x + (c / nrm) * SOFT_OFFSET * self.dx

    def _refine(self):
        """simulator.py:136-151: hard stays; soft spawns an extra point; inner/soft move to the
        exp(f)-weighted centroid of their components."""
        for p in self.cps.items():
            if p.kind == "hard":
                continue
            if p.kind == "soft":
                nx = self._soft_point(p)
                if nx is not None:
                    self.cps.insert(Item(nx, np.zeros(7)))
            self.cps.remove_near(p.x, EPS)
            w = np.array([np.exp(n.f) for n in p.near])
            xs = np.array([n.x for n in p.near])
            p.x = (w[:, None] * xs).sum(axis=0) / w.sum()
            self.cps.insert(p)

    def _resample(self):
        """simulator.py:169-198: scatter (creating a control point where none is near) and gap-fill."""
        for nu in self.comps.items():
            near = self.cps.near(nu.x, self.dx + EPS)
            if not near:
                near = [self.cps.insert(Item(nu.x.copy(), np.zeros(7)))]
            ws = [self.dx + 2.0 * EPS - float(np.linalg.norm(p.x - nu.x)) for p in near]
            sw = sum(ws)
            for p, w in zip(near, ws):
                p.f[nu.i] += (w / sw) * nu.f
        out = []
        for p in self.cps.items():
            f = p.f.copy()
            near = None
            for i in range(7):
                if f[i] < EPS:
                    if near is None:
                        near = self.cps.near(p.x, self.dx + EPS)
                    sw = sum(self.dx + 2.0 * EPS - float(np.linalg.norm(p.x - q.x)) for q in near if q.f[i] < EPS)
                    f[i] = (self.dx / sw) * self.f_rest[i]
            out.append(Item(p.x.copy(), f))
        self.moments_ = out
        self.comps.cells.clear()
