"""FP64 restatement of the reference GPU step (the oracle of the HIP engine).

Follows direbm/warp/physics.py:31-75, direbm/warp/propagation.py:21-246 and
direbm/warp/simulator.py:50-73.  The step is

    collide -> disperse -> cell-thinning -> refine -> resample (scatter + gap-fill/emit)

Arithmetic contract (what "parity" means):

* Every *discrete* decision is evaluated in the fp32 arithmetic the reference kernels define,
  with no fused multiply-add: component positions ``pos_m + C[k]*dx`` (propagation.py:34), the cell
  index ``floor(x / cs)`` with an IEEE fp32 division (propagation.py:48-49), the control-point
  centroid summed in fp32 in member order and divided by ``float32(n)`` (propagation.py:62-68),
  and the radius predicate ``sqrt(dx*dx + dy*dy) <= r`` (propagation.py:98, 175, 213).  Given the
  same fp32 input state, the oracle and the HIP engine therefore choose the same control points,
  in the same order, with the same members.
* Every *continuous* quantity (collision, exp-weighted centroids, scatter weights, gap-fill mass)
  is computed in FP64.

Extensions with no reference counterpart (parameters default to the reference behaviour):

* ``r_res``: resample radius (scatter and gap-fill); defaults to dx (propagation.py:175, 213).
* ``box``: periodic square box [0, L)^2 -- positions wrap after dispersion, cells wrap, distances
  use the minimum image.  A coordinate whose fp32 cell index rounds up to n (x within an ulp of
  L) is set to 0, so every point lies inside its own cell.
* ``two_level``: (cx, cy, r_in): components with |x - c|^2 < r_in^2 (fp32) are thinned on cells of
  size cs, the others on cells of size 2 cs.  Key = (coarse cell ix, coarse cell iy, sub) with
  sub = 2*(fx & 1) + (fy & 1) for fine cells and 4 for coarse cells; fx = floor(x / cs).
* ``circle``: (cx, cy, R), a circular obstacle with fluid outside: the v1 solver's Circle, specular
  bounce and removal of moments inside after the step (direbm/reference/boundary.py:20-70,
  direbm/reference/simulator.py:79-80,90-117), carried into this step's stages: a component
  k != 0 whose fp32 position is inside is replaced by its bounce parts, appended after the
  regular components.
* ``kappa_soft``: the per-object solver's soft_outer spawn (direbm/reference/simulator.py:134-166)
  carried into this step, which otherwise treats soft_outer points as inner (the reference GPU
  step does, direbm/warp/propagation.py:108).  A control point with kappa_hard < kappa <=
  kappa_soft spawns one extra point at its pre-refine position + n SOFT_OFFSET dx, n the unit
  vector of the sum of the directions c_k of its components within dx (counted with multiplicity,
  the components of kappa); no spawn when that sum is zero (the reference's |sum| <= 1e-8 -- on
  the lattice |sum| is 0 or >= 1).  The spawn position is computed in FP64 from the integer
  direction counts and rounded to fp32 (wrapped in a periodic box).  Spawned points are not
  refined (the reference refines a snapshot, simulator.py:136); they follow the regular refined
  points, in the order of their parents, and take part in the resample like any control point.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

from .constants import ALPHA, C2, DX, EPS_EMPTY, KAPPA_HARD, W2
from .physics import collide

F32 = np.float32
C2_32 = C2.astype(F32)
W2_32 = W2.astype(F32)
KEY_BIAS = 16384  # direbm/warp/propagation.py:17: valid for |floor(x/cs)| < 16384
SOFT_OFFSET = 2.0 * (1.0 - np.sqrt(3.0) / 2.0)  # direbm/reference/simulator.py:25-27


@dataclass
class Params:
    """Step parameters (direbm/warp/simulator.py:26-47 plus the extensions above)."""

    tau: float
    dx: float = DX
    alpha: float = ALPHA
    kappa_hard: int = KAPPA_HARD
    rho_rest: float = 1.0
    eps: float = EPS_EMPTY
    r_res: float | None = None
    box: float | None = None
    two_level: tuple | None = None
    circle: tuple | None = None
    kappa_soft: int = 0  # > kappa_hard: soft_outer spawn (0: off, the reference GPU step)

    @property
    def cs(self) -> float:
        return float(F32(self.dx / self.alpha))  # simulator.py:44, uploaded as fp32

    @property
    def rr(self) -> float:
        return float(self.dx if self.r_res is None else self.r_res)

    @property
    def ncell(self) -> int | None:
        """Fine cells per period (periodic box only); L / cs must be an integer."""
        if self.box is None:
            return None
        n = int(round(self.box / self.cs))
        if abs(n * self.cs - self.box) > 1e-6 * self.box:
            raise ValueError("periodic box must be a whole number of thinning cells")
        if self.two_level is not None and n % 2:
            raise ValueError("two-level thinning in a periodic box needs an even cell count")
        return n

    def f_rest(self):
        return self.rho_rest * W2  # simulator.py:47 (u_rest = 0)


# ----------------------------------------------------------------------------------- fp32 helpers
def wrap32(x, L, cs, n):
    """Wrap fp32 coordinates into [0, L) so that floor(x / cs) < n (see module docstring)."""
    x = np.asarray(x, dtype=F32).copy()
    L32, cs32 = F32(L), F32(cs)
    lo = x < 0
    x[lo] = x[lo] + L32
    hi = x >= L32
    x[hi] = x[hi] - L32
    bad = np.floor(x / cs32) >= n
    x[bad] = F32(0.0)
    return x


def min_image32(d, L):
    if L is None:
        return d
    L32 = F32(L)
    h = F32(0.5) * L32
    d = d.copy()
    d[d > h] -= L32
    d[d < -h] += L32
    return d


def dist32(ax, ay, bx, by, L=None):
    """fp32 distance exactly as the reference evaluates wp.length(a - b) (no FMA)."""
    ddx = min_image32((np.asarray(ax, F32) - np.asarray(bx, F32)).astype(F32), L)
    ddy = min_image32((np.asarray(ay, F32) - np.asarray(by, F32)).astype(F32), L)
    return np.sqrt((ddx * ddx + ddy * ddy).astype(F32)).astype(F32)


def delta64(a, b, L=None):
    """FP64 a - b of fp32 coordinates with the fp32 minimum-image convention."""
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    if L is not None:
        L64 = float(F32(L))
        d = np.where(d > 0.5 * L64, d - L64, d)
        d = np.where(d < -0.5 * L64, d + L64, d)
    return d


def pairs_within(a, b, r, L=None):
    """All (i, j) with dist32(a_i, b_j) <= r.  Returns (i, j, d32)."""
    a = np.asarray(a, F32)
    b = np.asarray(b, F32)
    if len(a) == 0 or len(b) == 0:
        z = np.zeros(0, np.int64)
        return z, z, np.zeros(0, F32)
    kw = {}
    if L is not None:
        kw["boxsize"] = float(F32(L)) * (1 + 1e-12)
    ta = cKDTree(a.astype(np.float64), **kw)
    tb = cKDTree(b.astype(np.float64), **kw)
    sm = ta.sparse_distance_matrix(tb, r * (1 + 1e-4) + 1e-6, output_type="ndarray")
    i = sm["i"].astype(np.int64)
    j = sm["j"].astype(np.int64)
    d = dist32(a[i, 0], a[i, 1], b[j, 0], b[j, 1], L)
    keep = d <= F32(r)
    return i[keep], j[keep], d[keep]


# --------------------------------------------------------------------------------- the obstacle
def _circle32(circle):
    c = np.asarray(circle[:2], F32)
    R = F32(circle[2])
    return c, R, F32(R * R)


def inside32(p, circle):
    """Strictly inside the circle, in fp32: (dx^2 + dy^2) < fp32(R * R)
    (direbm/reference/boundary.py:24-26 with the radius predicate's rounding contract)."""
    c, _, R2 = _circle32(circle)
    p = np.asarray(p, F32).reshape(-1, 2)
    dx = (p[:, 0] - c[0]).astype(F32)
    dy = (p[:, 1] - c[1]).astype(F32)
    return ((dx * dx).astype(F32) + (dy * dy).astype(F32)).astype(F32) < R2


def ray_hit64(x0, x1, circle):
    """First crossing of the segment x0 -> x1 with the circle and the outward normal there, in
    FP64 (direbm/reference/boundary.py:28-47).  Returns (hit, normal)."""
    c, R, _ = _circle32(circle)
    c = c.astype(np.float64)
    x0 = np.asarray(x0, np.float64)
    d = np.asarray(x1, np.float64) - x0
    f = x0 - c
    a = d[0] * d[0] + d[1] * d[1]
    b = 2.0 * (f[0] * d[0] + f[1] * d[1])
    cc = (f[0] * f[0] + f[1] * f[1]) - float(R) * float(R)
    disc = b * b - 4.0 * a * cc
    if a == 0.0 or disc < 0.0:
        return False, None
    sq = np.sqrt(disc)
    for t in ((-b - sq) / (2.0 * a), (-b + sq) / (2.0 * a)):
        if 0.0 <= t <= 1.0:
            p = x0 + t * d
            return True, (p - c) / float(R)
    return False, None


def reflect64(c, n):
    """c - 2 (c.n) n (direbm/reference/boundary.py:56-60)."""
    c = np.asarray(c, np.float64)
    n = np.asarray(n, np.float64)
    return c - 2.0 * (c[0] * n[0] + c[1] * n[1]) * n


def split_hex(d):
    """The two D2Q7 directions adjacent to d by angle, with linear weights (1 - frac, frac)
    (direbm/reference/boundary.py:63-70).  Returns [(j, w), (j', w')]."""
    theta = np.arctan2(d[1], d[0]) % (2.0 * np.pi)
    seg = theta / (np.pi / 3.0)
    k = int(np.floor(seg))
    frac = seg - k
    return [(k % 6 + 1, 1.0 - frac), ((k + 1) % 6 + 1, frac)]


# ------------------------------------------------------------------------------------ the stages
def disperse(pos_m, f_m, dx, prm: Params | None = None):
    """Moment -> 7 components.  direbm/warp/propagation.py:21-36.

    Component c = 7*i + k sits at fp32 ``pos_m[i] + C[k]*dx`` (product rounded, then sum
    rounded), carries f_m[i, k] and direction k.  Returns (pos_c fp32 (7M,2), val_c, dir_c).
    """
    pos_m = np.asarray(pos_m, F32).reshape(-1, 2)
    f_m = np.asarray(f_m)
    step = (C2_32 * F32(dx)).astype(F32)
    pc = (pos_m[:, None, :] + step[None, :, :]).astype(F32).reshape(-1, 2)
    if prm is not None and prm.box is not None:
        n = prm.ncell
        pc = np.stack([wrap32(pc[:, 0], prm.box, prm.cs, n), wrap32(pc[:, 1], prm.box, prm.cs, n)], 1)
    val = f_m.reshape(-1)
    dirs = np.tile(np.arange(7, dtype=np.int32), len(pos_m))
    if prm is not None and prm.circle is not None and len(pc):
        # Bounce (direbm/reference/simulator.py:90-117): the ray from the moment to the component
        # is hit-tested, c_k reflected about the normal and split; each part sits at fp32(pos +
        # fp32(C_j dx)), or at the moment when that is inside too, with value f_k w_j.  Parts come
        # after every regular component, in (moment, direction, part) order.  (The circle keeps dx
        # clear of a periodic box's edges, so wrapped and unwrapped inside tests agree.)
        hit = inside32(pc, prm.circle) & (dirs != 0)
        extra_p, extra_v, extra_d = [], [], []
        for c in np.flatnonzero(hit):
            i, k = divmod(int(c), 7)
            ok, nrm = ray_hit64(pos_m[i], pc[c], prm.circle)
            if not ok:
                continue
            for j, w in split_hex(reflect64(C2[k], nrm)):
                q = (pos_m[i] + step[j]).astype(F32)
                if prm.box is not None:
                    q = np.array([wrap32(q[0:1], prm.box, prm.cs, prm.ncell)[0],
                                  wrap32(q[1:2], prm.box, prm.cs, prm.ncell)[0]], F32)
                if inside32(q, prm.circle)[0]:
                    q = pos_m[i].copy()
                extra_p.append(q)
                extra_v.append(float(val[c]) * w)
                extra_d.append(j)
        keep = ~hit
        pc = np.concatenate([pc[keep], np.asarray(extra_p, F32).reshape(-1, 2)])
        val = np.concatenate([np.asarray(val, np.float64)[keep], np.asarray(extra_v, np.float64)])
        dirs = np.concatenate([dirs[keep], np.asarray(extra_d, np.int32)])
    return pc, val, dirs


def fine_cells(pos, prm: Params):
    """fx = floor(x / cs) with an fp32 division (propagation.py:48-49), wrapped if periodic."""
    pos = np.asarray(pos, F32)
    cs32 = F32(prm.cs)
    f = np.floor((pos / cs32).astype(F32)).astype(np.int64)
    if prm.box is not None:
        f %= prm.ncell
    elif np.abs(f).max(initial=0) >= KEY_BIAS:
        raise OverflowError("cell index outside the reference key range |floor(x/cs)| < 16384")
    return f[:, 0], f[:, 1]


def thin_keys(pos_c, prm: Params):
    """(qx, qy, sub) per component; lexicographic order == the reference key order."""
    fx, fy = fine_cells(pos_c, prm)
    if prm.two_level is None:
        return fx, fy, np.zeros_like(fx)
    cx, cy, rin = (F32(v) for v in prm.two_level)
    pos = np.asarray(pos_c, F32)
    ex = (pos[:, 0] - cx).astype(F32)
    ey = (pos[:, 1] - cy).astype(F32)
    inner = (ex * ex + ey * ey).astype(F32) < (rin * rin).astype(F32)
    sub = np.where(inner, ((fx & 1) << 1) | (fy & 1), 4)
    return fx >> 1, fy >> 1, sub


@dataclass
class Thinning:
    order: np.ndarray  # stable permutation of components into key order
    start: np.ndarray  # run start (into order) per control point
    length: np.ndarray  # run length per control point
    key: np.ndarray  # (P, 3) (qx, qy, sub) per control point
    cp32: np.ndarray  # (P, 2) fp32 centroid, bit-exact reference semantics
    cp64: np.ndarray  # (P, 2) FP64 centroid


def thin(pos_c, prm: Params) -> Thinning:
    """Cell-thinning control points.  direbm/warp/propagation.py:39-68, 130-153.

    Stable sort by cell key (ties in ascending component index), one control point per run, in
    ascending key order, at the centroid of its members summed in fp32 in member order.
    """
    pos_c = np.asarray(pos_c, F32)
    qx, qy, sub = thin_keys(pos_c, prm)
    order = np.lexsort((sub, qy, qx))  # stable
    kq = np.stack([qx[order], qy[order], sub[order]], 1)
    n = len(order)
    if n == 0:
        e = np.zeros(0, np.int64)
        return Thinning(order, e, e, np.zeros((0, 3), np.int64), np.zeros((0, 2), F32), np.zeros((0, 2)))
    head = np.ones(n, bool)
    head[1:] = np.any(kq[1:] != kq[:-1], axis=1)
    start = np.flatnonzero(head)
    length = np.diff(np.append(start, n))
    ps = pos_c[order]
    acc = np.zeros((len(start), 2), F32)
    for j in range(int(length.max())):
        m = length > j
        acc[m] = (acc[m] + ps[start[m] + j]).astype(F32)
    cp32 = (acc / length.astype(F32)[:, None]).astype(F32)
    run = np.cumsum(head) - 1
    s64 = np.zeros((len(start), 2))
    np.add.at(s64, run, ps.astype(np.float64))
    cp64 = s64 / length[:, None]
    return Thinning(order, start, length, kq[start], cp32, cp64)


def refine(cp, pos_c, val_c, dir_c, dx, kappa_hard, prm: Params | None = None):
    """Refine control points.  direbm/warp/propagation.py:77-112, 249-275.

    Over components with fp32 distance <= dx: kappa = number of distinct directions; kappa <=
    kappa_hard keeps the point, otherwise it moves to the exp(val)-weighted centroid
    sum(w q) / sum(w) (periodic box: p + sum(w (q - p)) / sum(w), wrapped).  Returns
    (new positions FP64 (P,2), kappa (P,)).
    """
    cp = np.asarray(cp, F32).reshape(-1, 2)
    pos_c = np.asarray(pos_c, F32).reshape(-1, 2)
    L = None if prm is None else prm.box
    ri, ci, _ = pairs_within(cp, pos_c, dx, L)
    P = len(cp)
    mask = np.zeros(P, np.int64)
    np.bitwise_or.at(mask, ri, np.left_shift(1, np.asarray(dir_c, np.int64)[ci]))
    kappa = np.array([bin(int(m)).count("1") for m in mask], np.int64) if P else np.zeros(0, np.int64)
    w = np.exp(np.asarray(val_c, np.float64)[ci])
    sw = np.bincount(ri, w, minlength=P)
    new = cp.astype(np.float64).copy()
    inner = (kappa > int(kappa_hard)) & (sw > 0)
    if L is None:
        sx = np.bincount(ri, w * pos_c[ci, 0], minlength=P)
        sy = np.bincount(ri, w * pos_c[ci, 1], minlength=P)
        new[inner, 0] = sx[inner] / sw[inner]
        new[inner, 1] = sy[inner] / sw[inner]
    else:
        sx = np.bincount(ri, w * delta64(pos_c[ci, 0], cp[ri, 0], L), minlength=P)
        sy = np.bincount(ri, w * delta64(pos_c[ci, 1], cp[ri, 1], L), minlength=P)
        L64 = float(F32(L))
        new[inner, 0] = np.mod(cp[inner, 0] + sx[inner] / sw[inner], L64)
        new[inner, 1] = np.mod(cp[inner, 1] + sy[inner] / sw[inner], L64)
    return new, kappa


def spawn(cp, pos_c, dir_c, kappa, dx, kappa_hard, kappa_soft, prm: Params | None = None):
    """Soft_outer spawn positions (direbm/reference/simulator.py:141-166): for each control point
    with kappa_hard < kappa <= kappa_soft, the sum of the directions of its components within dx
    (as integer counts: 2 sum c_x and the signed count of the c_y = +-sqrt(3)/2 directions), and,
    when it is nonzero, cp + n SOFT_OFFSET dx in FP64 rounded to fp32.  Returns (parent indices,
    spawn positions (S, 2) fp32), parents ascending."""
    cp = np.asarray(cp, F32).reshape(-1, 2)
    pos_c = np.asarray(pos_c, F32).reshape(-1, 2)
    dirs = np.asarray(dir_c, np.int64)
    L = None if prm is None else prm.box
    soft = (np.asarray(kappa) > int(kappa_hard)) & (np.asarray(kappa) <= int(kappa_soft))
    if not soft.any():
        return np.zeros(0, np.int64), np.zeros((0, 2), F32)
    ri, ci, _ = pairs_within(cp, pos_c, dx, L)
    keep = soft[ri]
    ri, k = ri[keep], dirs[ci[keep]]
    ix2 = np.array([0, 2, 1, -1, -2, -1, 1], np.int64)  # 2 c_x
    iy = np.array([0, 0, 1, 1, 0, -1, -1], np.int64)  # c_y / (sqrt(3) / 2)
    P = len(cp)
    sx2 = np.bincount(ri, ix2[k], minlength=P)
    sy = np.bincount(ri, iy[k], minlength=P)
    par = np.flatnonzero(soft & ((sx2 != 0) | (sy != 0)))
    cx = 0.5 * sx2[par].astype(np.float64)
    cy = (np.sqrt(3.0) / 2.0) * sy[par].astype(np.float64)
    nrm = np.sqrt(cx * cx + cy * cy)
    off = SOFT_OFFSET * float(F32(dx))  # dx as the device holds it
    out = np.stack([cp[par, 0].astype(np.float64) + cx / nrm * off, cp[par, 1].astype(np.float64) + cy / nrm * off],
                   1).astype(F32)
    if L is not None:
        n = prm.ncell
        out = np.stack([wrap32(out[:, 0], L, prm.cs, n), wrap32(out[:, 1], L, prm.cs, n)], 1)
    return par, out


@dataclass
class Resampled:
    pos: np.ndarray  # (P, 2) emitted moment positions (= control points)
    f: np.ndarray  # (P, 7) FP64 emitted distributions
    cp_f: np.ndarray  # (P, 7) FP64 scatter result before gap-fill
    sw: np.ndarray  # (nc,) FP64 per-component scatter normaliser
    dropped: float  # component mass with no control point in range
    injected: float  # mass added by gap-fill


def resample(pos_c, val_c, dir_c, cp, dx, f_rest, prm: Params | None = None) -> Resampled:
    """Scatter + gap-fill/emit.  direbm/warp/propagation.py:156-218, 221-246.

    Scatter: each component c with sw_c = sum_j (r - d_cj) over control points within r
    delivers (r - d_cj) / sw_c * val_c into cp_f[j, dir_c]; sw_c <= 0 drops it.
    Gap-fill: a direction with cp_f[p, i] < eps becomes (r / sw) f_rest[i] with
    sw = sum (r - d) over control points within r that are also empty in i (self included).
    """
    cp64 = np.asarray(cp, np.float64).reshape(-1, 2)  # distances (== cp for fp32 input)
    cp = cp64.astype(F32)  # radius predicates
    pos_c = np.asarray(pos_c, F32).reshape(-1, 2)
    val = np.asarray(val_c, np.float64).reshape(-1)
    dirs = np.asarray(dir_c, np.int64).reshape(-1)
    L = None if prm is None else prm.box
    r = float(dx if prm is None else prm.rr)
    eps = EPS_EMPTY if prm is None else prm.eps
    P, nc = len(cp), len(pos_c)
    ci, pj, _ = pairs_within(pos_c, cp, r, L)
    d64 = np.hypot(delta64(cp64[pj, 0], pos_c[ci, 0], L), delta64(cp64[pj, 1], pos_c[ci, 1], L))
    w = r - d64
    sw = np.bincount(ci, w, minlength=nc)
    alive = sw > 0
    cp_f = np.zeros((P, 7))
    ok = alive[ci]
    np.add.at(cp_f, (pj[ok], dirs[ci[ok]]), w[ok] / sw[ci[ok]] * val[ci[ok]])
    dropped = float(val[~alive].sum()) if nc else 0.0

    empty = cp_f < eps
    out = cp_f.copy()
    if P:
        ri, rj, _ = pairs_within(cp, cp, r, L)
        dd = np.hypot(delta64(cp64[rj, 0], cp64[ri, 0], L), delta64(cp64[rj, 1], cp64[ri, 1], L))
        wr = r - dd
        fr = np.asarray(f_rest, np.float64)
        for i in range(7):
            swi = np.bincount(ri, wr * empty[rj, i], minlength=P)
            fill = empty[:, i] & (swi > 0)
            out[fill, i] = (r / swi[fill]) * fr[i]
    injected = float((out - cp_f)[empty].sum())
    return Resampled(cp64, out, cp_f, sw, dropped, injected)


@dataclass
class StepResult:
    pos: np.ndarray  # (P, 2) FP64 new moment positions
    f: np.ndarray  # (P, 7) FP64 new distributions
    thin: Thinning
    refined: np.ndarray
    kappa: np.ndarray
    res: Resampled
    f_star: np.ndarray  # collided distributions of the input moments


def step(pos_m, f_m, prm: Params, round_positions: bool = True) -> StepResult:
    """One DiReBM step on an fp32 state.  direbm/warp/simulator.py:50-73.

    Positions travel between stages as fp32 (what the device stores); values stay FP64.
    round_positions=False keeps the refined positions in FP64 for the distances of the resample
    (radius predicates stay fp32): the smooth form central differences of the adjoint need.
    """
    pos_m = np.asarray(pos_m, F32).reshape(-1, 2)
    f_star = collide(np.asarray(f_m, np.float64).reshape(-1, 7), prm.tau)
    pos_c, val_c, dir_c = disperse(pos_m, f_star, prm.dx, prm)
    th = thin(pos_c, prm)
    refined, kappa = refine(th.cp32, pos_c, val_c, dir_c, prm.dx, prm.kappa_hard, prm)
    cp_next = refined.astype(F32)
    if prm.box is not None:
        n = prm.ncell
        cp_next = np.stack([wrap32(cp_next[:, 0], prm.box, prm.cs, n), wrap32(cp_next[:, 1], prm.box, prm.cs, n)], 1)
    if prm.kappa_soft > prm.kappa_hard:
        _, born = spawn(th.cp32, pos_c, dir_c, kappa, prm.dx, prm.kappa_hard, prm.kappa_soft, prm)
        cp_next = np.concatenate([cp_next, born]) if len(born) else cp_next
        refined = np.concatenate([refined, born.astype(np.float64)]) if len(born) else refined
    if not round_positions:
        cp_next = refined
    res = resample(pos_c, val_c, dir_c, cp_next, prm.dx, prm.f_rest(), prm)
    pos_out, f_out = cp_next.astype(np.float64), res.f
    if prm.circle is not None and len(cp_next):
        # moments inside the obstacle are removed after resampling (simulator.py:79-80)
        out = ~inside32(np.asarray(cp_next, F32), prm.circle)
        pos_out, f_out = pos_out[out], f_out[out]
    return StepResult(pos_out, f_out, th, refined, kappa, res, f_star)


class Simulator:
    """Oracle twin of GpuSimulator (direbm/warp/simulator.py:25-77): state kept as fp32."""

    def __init__(self, pos_m, f_m, tau, *, dx=DX, alpha=ALPHA, kappa_hard=KAPPA_HARD, rho_rest=1.0,
                 r_res=None, box=None, two_level=None, circle=None, kappa_soft=0):
        self.prm = Params(tau=float(tau), dx=float(dx), alpha=float(alpha), kappa_hard=int(kappa_hard),
                          rho_rest=float(rho_rest), r_res=r_res, box=box, two_level=two_level, circle=circle,
                          kappa_soft=int(kappa_soft))
        self.pos_m = np.asarray(pos_m, F32).reshape(-1, 2).copy()
        self.f_m = np.asarray(f_m, F32).reshape(-1, 7).copy()
        self.iteration = 0
        self.last: StepResult | None = None

    def step(self):
        r = step(self.pos_m, self.f_m, self.prm)
        self.pos_m = r.pos.astype(F32)
        self.f_m = r.f.astype(F32)
        self.last = r
        self.iteration += 1

    def moments(self):
        return self.pos_m.copy(), self.f_m.copy()
