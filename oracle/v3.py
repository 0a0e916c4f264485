"""FP64 restatement of the DiReBM step in 3D on the D3Q13 icosahedral lattice (test
infrastructure: the checker of the 3D HIP engine).

The reference runs 3D only in its per-object solver (direbm/reference/simulator.py:30-198 with
lattice=D3Q13, direbm/lattices.py:44-61); its notes plan the GPU step for 3D as the 2D step with
vec3 positions, 13 directions and cs^2 = 1/5 (HANDOFF.md:135-172).  This is that step, with the
same arithmetic contract as oracle/v2.py:

* discrete decisions in fp32 with one rounding per operation: component position
  fp32(pos + fp32(C_k dx)), cell index floor(fp32(x / cs)) per axis, radius predicate
  sqrt_rn(fp32(fp32(dx^2) + fp32(dy^2)) + fp32(dz^2)) <= r, centroid summed in fp32 in member
  order and divided by float32(n);
* continuous quantities in FP64.

Cell keys are (ix, iy, iz) lexicographic (x-major), the 3D extension of the reference's
ix * 32768 + iy (direbm/warp/propagation.py:46-50).  Options: r_res (resample radius) and box
(periodic cube [0, L)^3), as in oracle/v2.py.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

from .constants import ALPHA, C3, CS2_3D, DX, EPS_EMPTY, KAPPA_HARD, W3

F32 = np.float32
C3_32 = C3.astype(F32)
W3_32 = W3.astype(F32)


@dataclass
class Params:
    tau: float
    dx: float = DX
    alpha: float = ALPHA
    kappa_hard: int = KAPPA_HARD
    rho_rest: float = 1.0
    eps: float = EPS_EMPTY
    r_res: float | None = None
    box: float | None = None
    sphere: tuple | None = None  # (cx, cy, cz, R): spherical obstacle, fluid outside

    @property
    def cs(self) -> float:
        return float(F32(self.dx / self.alpha))

    @property
    def rr(self) -> float:
        return float(self.dx if self.r_res is None else self.r_res)

    @property
    def ncell(self) -> int | None:
        if self.box is None:
            return None
        n = int(round(self.box / self.cs))
        if abs(n * self.cs - self.box) > 1e-6 * self.box:
            raise ValueError("periodic box must be a whole number of thinning cells")
        return n

    def f_rest(self):
        return self.rho_rest * W3


def equilibrium(rho, u):
    """f_eq_i = rho W_i (1 + cu/cs2 + cu^2/(2 cs2^2) - u.u/(2 cs2)) on D3Q13 (direbm/physics.py:29-45)."""
    rho = np.asarray(rho, np.float64)
    u = np.asarray(u, np.float64)
    inv = 1.0 / CS2_3D
    cu = u @ C3.T
    usq = (u * u).sum(-1)
    return rho[..., None] * W3 * (1.0 + inv * cu + 0.5 * inv * inv * cu * cu - 0.5 * inv * usq[..., None])


def collide(f, tau):
    """BGK (direbm/physics.py:48-55) with the rho == 0 guard (u = 0)."""
    f = np.asarray(f, np.float64)
    rho = f.sum(-1)
    j = f @ C3
    safe = np.where(rho == 0.0, 1.0, rho)
    u = j / safe[..., None]
    return f + (equilibrium(rho, u) - f) / tau


def wrap32(x, L, cs, n):
    x = np.asarray(x, dtype=F32).copy()
    L32 = F32(L)
    lo = x < 0
    x[lo] = x[lo] + L32
    hi = x >= L32
    x[hi] = x[hi] - L32
    bad = np.floor(x / F32(cs)) >= n
    x[bad] = F32(0.0)
    return x


def _mi32(d, L):
    if L is None:
        return d
    L32 = F32(L)
    h = F32(0.5) * L32
    d = d.copy()
    d[d > h] -= L32
    d[d < -h] += L32
    return d


def dist32(a, b, L=None):
    """fp32 distance as the kernels evaluate it: ((dx^2 + dy^2) + dz^2), then sqrt, no FMA."""
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    d = [_mi32((a[:, k] - b[:, k]).astype(F32), L) for k in range(3)]
    s = ((d[0] * d[0]).astype(F32) + (d[1] * d[1]).astype(F32)).astype(F32)
    s = (s + (d[2] * d[2]).astype(F32)).astype(F32)
    return np.sqrt(s).astype(F32)


def delta64(a, b, L=None):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    if L is not None:
        L64 = float(F32(L))
        d = np.where(d > 0.5 * L64, d - L64, d)
        d = np.where(d < -0.5 * L64, d + L64, d)
    return d


def pairs_within(a, b, r, L=None):
    a = np.asarray(a, F32).reshape(-1, 3)
    b = np.asarray(b, F32).reshape(-1, 3)
    if len(a) == 0 or len(b) == 0:
        z = np.zeros(0, np.int64)
        return z, z
    kw = {}
    if L is not None:
        kw["boxsize"] = float(F32(L)) * (1 + 1e-12)
    ta = cKDTree(a.astype(np.float64), **kw)
    tb = cKDTree(b.astype(np.float64), **kw)
    sm = ta.sparse_distance_matrix(tb, r * (1 + 1e-4) + 1e-6, output_type="ndarray")
    i = sm["i"].astype(np.int64)
    j = sm["j"].astype(np.int64)
    keep = dist32(a[i], b[j], L) <= F32(r)
    return i[keep], j[keep]


def _sphere32(sph):
    c = np.asarray(sph[:3], F32)
    R = F32(sph[3])
    return c, R, F32(R * R)


def inside32(p, sph):
    """Strictly inside the sphere, in fp32: ((dx^2 + dy^2) + dz^2) < fp32(R * R)
    (direbm/reference/boundary.py:24-26 with the radius predicate's rounding contract)."""
    c, _, R2 = _sphere32(sph)
    p = np.asarray(p, F32).reshape(-1, 3)
    d = [(p[:, k] - c[k]).astype(F32) for k in range(3)]
    s = ((d[0] * d[0]).astype(F32) + (d[1] * d[1]).astype(F32)).astype(F32)
    s = (s + (d[2] * d[2]).astype(F32)).astype(F32)
    return s < R2


def ray_hit64(x0, x1, sph):
    """First crossing of the segment x0 -> x1 with the sphere and the outward normal there, in FP64
    (direbm/reference/boundary.py:28-47).  Returns (hit, normal)."""
    c, R, _ = _sphere32(sph)
    x0 = np.asarray(x0, np.float64)
    d = np.asarray(x1, np.float64) - x0
    f = x0 - c.astype(np.float64)
    a = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    b = 2.0 * ((f[0] * d[0] + f[1] * d[1]) + f[2] * d[2])
    cc = ((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]) - float(R) * float(R)
    disc = b * b - 4.0 * a * cc
    if a == 0.0 or disc < 0.0:
        return False, None
    sq = np.sqrt(disc)
    for t in ((-b - sq) / (2.0 * a), (-b + sq) / (2.0 * a)):
        if 0.0 <= t <= 1.0:
            p = x0 + t * d
            return True, (p - c.astype(np.float64)) / float(R)
    return False, None


def reflect64(c, n):
    """c - 2 (c.n) n (direbm/reference/boundary.py:56-60), dot product ((c0 n0 + c1 n1) + c2 n2)."""
    c = np.asarray(c, np.float64)
    n = np.asarray(n, np.float64)
    cn = (c[0] * n[0] + c[1] * n[1]) + c[2] * n[2]
    return c - 2.0 * cn * n


def split_nd(d):
    """Split a direction among the twelve non-rest directions with weights max(d.c_j, 0)^2 / sum,
    parts above 1e-9 kept (direbm/reference/boundary.py:73-85).  Dot products
    ((c0 d0 + c1 d1) + c2 d2) and the weight sum over j = 1..12 in sequence, in FP64.
    Returns [(j, w)]."""
    d = np.asarray(d, np.float64)
    dots = [(C3[j, 0] * d[0] + C3[j, 1] * d[1]) + C3[j, 2] * d[2] for j in range(1, 13)]
    w = [max(x, 0.0) * max(x, 0.0) for x in dots]
    s = 0.0
    for x in w:
        s += x
    if s <= 0.0:
        return [(int(np.argmax(dots)) + 1, 1.0)]
    return [(j + 1, x / s) for j, x in enumerate(w) if x / s > 1e-9]


def disperse(pos_m, f_m, prm: Params):
    """Moment -> 13 components at fp32(pos + fp32(C_k dx)) (wrapped in a periodic box).

    With a sphere: a component k != 0 whose position is inside bounces (direbm/reference/
    simulator.py:90-117): the ray from its moment to it is hit-tested, c_k reflected and split into
    parts at fp32(pos + fp32(C_j dx)) (the moment's position when that is inside too), each of value
    f_k w_j.  The regular components come first in (moment, direction) order, then every part in
    (moment, direction, j) order."""
    pos_m = np.asarray(pos_m, F32).reshape(-1, 3)
    step = (C3_32 * F32(prm.dx)).astype(F32)
    pc = (pos_m[:, None, :] + step[None, :, :]).astype(F32).reshape(-1, 3)
    val = np.asarray(f_m).reshape(-1)
    dirs = np.tile(np.arange(13, dtype=np.int32), len(pos_m))
    if prm.sphere is not None and len(pc):
        hit = inside32(pc, prm.sphere) & (dirs != 0)
        extra_p, extra_v, extra_d = [], [], []
        for c in np.flatnonzero(hit):
            i, k = divmod(int(c), 13)
            ok, n = ray_hit64(pos_m[i], pc[c], prm.sphere)
            if not ok:
                continue
            for j, w in split_nd(reflect64(C3[k], n)):
                q = (pos_m[i] + step[j]).astype(F32)
                if inside32(q, prm.sphere)[0]:
                    q = pos_m[i].copy()
                extra_p.append(q)
                extra_v.append(float(val[c]) * w)
                extra_d.append(j)
        keep = ~hit
        pc = np.concatenate([pc[keep], np.asarray(extra_p, F32).reshape(-1, 3)])
        val = np.concatenate([val[keep], np.asarray(extra_v, np.float64)])
        dirs = np.concatenate([dirs[keep], np.asarray(extra_d, np.int32)])
    if prm.box is not None:
        n = prm.ncell
        pc = np.stack([wrap32(pc[:, k], prm.box, prm.cs, n) for k in range(3)], 1)
    return pc, val, dirs


def cells(pos, prm: Params):
    f = np.floor((np.asarray(pos, F32) / F32(prm.cs)).astype(F32)).astype(np.int64)
    if prm.box is not None:
        f %= prm.ncell
    return f


@dataclass
class Thinning:
    order: np.ndarray
    start: np.ndarray
    length: np.ndarray
    key: np.ndarray
    cp32: np.ndarray


def thin(pos_c, prm: Params) -> Thinning:
    """Stable sort by (ix, iy, iz), one control point per occupied cell at the fp32 centroid of its
    members summed in member order (direbm/warp/propagation.py:54-68,130-153 in 3D)."""
    pos_c = np.asarray(pos_c, F32)
    k = cells(pos_c, prm)
    order = np.lexsort((k[:, 2], k[:, 1], k[:, 0]))
    ks = k[order]
    n = len(order)
    if n == 0:
        e = np.zeros(0, np.int64)
        return Thinning(order, e, e, np.zeros((0, 3), np.int64), np.zeros((0, 3), F32))
    head = np.ones(n, bool)
    head[1:] = np.any(ks[1:] != ks[:-1], axis=1)
    start = np.flatnonzero(head)
    length = np.diff(np.append(start, n))
    ps = pos_c[order]
    acc = np.zeros((len(start), 3), F32)
    for j in range(int(length.max())):
        m = length > j
        acc[m] = (acc[m] + ps[start[m] + j]).astype(F32)
    cp32 = (acc / length.astype(F32)[:, None]).astype(F32)
    return Thinning(order, start, length, ks[start], cp32)


def refine(cp, pos_c, val_c, dir_c, prm: Params):
    """kappa = distinct directions within dx; kappa <= kappa_hard keeps, else the exp(val)-weighted
    centroid (periodic: p + sum w (q - p) / sum w, wrapped).  Returns (FP64 positions, kappa)."""
    cp = np.asarray(cp, F32).reshape(-1, 3)
    pos_c = np.asarray(pos_c, F32).reshape(-1, 3)
    L = prm.box
    ri, ci = pairs_within(cp, pos_c, prm.dx, L)
    P = len(cp)
    mask = np.zeros(P, np.int64)
    np.bitwise_or.at(mask, ri, np.left_shift(1, np.asarray(dir_c, np.int64)[ci]))
    kappa = np.array([bin(int(m)).count("1") for m in mask], np.int64) if P else np.zeros(0, np.int64)
    w = np.exp(np.asarray(val_c, np.float64)[ci])
    sw = np.bincount(ri, w, minlength=P)
    new = cp.astype(np.float64).copy()
    inner = (kappa > prm.kappa_hard) & (sw > 0)
    for k in range(3):
        if L is None:
            s = np.bincount(ri, w * pos_c[ci, k], minlength=P)
            new[inner, k] = s[inner] / sw[inner]
        else:
            s = np.bincount(ri, w * delta64(pos_c[ci, k], cp[ri, k], L), minlength=P)
            new[inner, k] = np.mod(cp[inner, k] + s[inner] / sw[inner], float(F32(L)))
    return new, kappa


@dataclass
class Resampled:
    pos: np.ndarray
    f: np.ndarray
    cp_f: np.ndarray
    dropped: float
    injected: float


def resample(pos_c, val_c, dir_c, cp, prm: Params) -> Resampled:
    """Scatter with weights (R - d) / sum and gap-fill of empty directions (< eps) with
    (R / sw) f_rest over also-empty control points within R (direbm/warp/propagation.py:156-246)."""
    cp = np.asarray(cp, F32).reshape(-1, 3)
    pos_c = np.asarray(pos_c, F32).reshape(-1, 3)
    val = np.asarray(val_c, np.float64).reshape(-1)
    dirs = np.asarray(dir_c, np.int64).reshape(-1)
    L, r = prm.box, prm.rr
    P, nc = len(cp), len(pos_c)
    ci, pj = pairs_within(pos_c, cp, r, L)
    d = np.sqrt(sum(delta64(cp[pj, k], pos_c[ci, k], L) ** 2 for k in range(3)))
    w = r - d
    sw = np.bincount(ci, w, minlength=nc)
    alive = sw > 0
    cp_f = np.zeros((P, 13))
    ok = alive[ci]
    np.add.at(cp_f, (pj[ok], dirs[ci[ok]]), w[ok] / sw[ci[ok]] * val[ci[ok]])
    dropped = float(val[~alive].sum()) if nc else 0.0
    empty = cp_f < prm.eps
    out = cp_f.copy()
    if P:
        ri, rj = pairs_within(cp, cp, r, L)
        dd = np.sqrt(sum(delta64(cp[rj, k], cp[ri, k], L) ** 2 for k in range(3)))
        wr = r - dd
        fr = prm.f_rest()
        for i in range(13):
            swi = np.bincount(ri, wr * empty[rj, i], minlength=P)
            fill = empty[:, i] & (swi > 0)
            out[fill, i] = (r / swi[fill]) * fr[i]
    injected = float((out - cp_f)[empty].sum())
    return Resampled(cp.astype(np.float64), out, cp_f, dropped, injected)


@dataclass
class StepResult:
    pos: np.ndarray
    f: np.ndarray
    thin: Thinning
    kappa: np.ndarray
    res: Resampled


def step(pos_m, f_m, prm: Params) -> StepResult:
    """One 3D step on an fp32 state: collide, disperse, thin, refine, resample."""
    pos_m = np.asarray(pos_m, F32).reshape(-1, 3)
    f_star = collide(np.asarray(f_m, np.float64).reshape(-1, 13), prm.tau)
    pos_c, val_c, dir_c = disperse(pos_m, f_star, prm)
    th = thin(pos_c, prm)
    refined, kappa = refine(th.cp32, pos_c, val_c, dir_c, prm)
    cp_next = refined.astype(F32)
    if prm.box is not None:
        n = prm.ncell
        cp_next = np.stack([wrap32(cp_next[:, k], prm.box, prm.cs, n) for k in range(3)], 1)
    res = resample(pos_c, val_c, dir_c, cp_next, prm)
    if prm.sphere is not None:
        # control points inside the obstacle are dropped after resampling (simulator.py:79-80)
        keep = ~inside32(cp_next, prm.sphere)
        return StepResult(cp_next[keep].astype(np.float64), res.f[keep], th, kappa, res)
    return StepResult(cp_next.astype(np.float64), res.f, th, kappa, res)


def bin_fields3(x, f, xmin, xmax, h):
    """3D binning of mass and momentum (the 3D counterpart of direbm/fields.py:17-41, which the
    reference notes is 2D only): rho = cell mass / h^3, u = momentum / mass; arrays [iz, iy, ix]."""
    x = np.asarray(x, np.float64)
    f = np.asarray(f, np.float64)
    edges = np.arange(xmin, xmax + h, h)
    m, _ = np.histogramdd(x, bins=[edges] * 3, weights=f.sum(1))
    j = f @ C3
    p = [np.histogramdd(x, bins=[edges] * 3, weights=j[:, k])[0] for k in range(3)]
    m = m.transpose(2, 1, 0)
    p = [q.transpose(2, 1, 0) for q in p]
    safe = np.where(m == 0, 1.0, m)
    u = np.stack([q / safe for q in p], -1)
    u[m == 0] = 0.0
    return m / h**3, u
