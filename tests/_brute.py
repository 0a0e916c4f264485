"""Pure-Python loop restatements of the reference stages, for tiny inputs only.

They pin oracle/v2.py by a structurally different computation: explicit double loops over all
pairs (no k-d tree, no vectorisation, no sorting), with the predicates written from the
reference kernels (direbm/warp/propagation.py:88-112, 168-186, 203-218) and fp32 distances.
"""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def d32(ax, ay, bx, by):
    ddx = F32(F32(ax) - F32(bx))
    ddy = F32(F32(ay) - F32(by))
    return F32(np.sqrt(F32(F32(ddx * ddx) + F32(ddy * ddy))))


def thin_dict(pos_c, cs):
    """One control point per occupied cell (dict of cells, insertion = component order)."""
    cells = {}
    for c, (x, y) in enumerate(pos_c):
        key = (math.floor(F32(F32(x) / F32(cs))), math.floor(F32(F32(y) / F32(cs))))
        cells.setdefault(key, []).append(c)
    out = []
    for key in sorted(cells):
        members = cells[key]
        sx = sum(float(pos_c[c][0]) for c in members) / len(members)
        sy = sum(float(pos_c[c][1]) for c in members) / len(members)
        out.append((key, members, (sx, sy)))
    return out


def refine_loops(cp, pos_c, val_c, dir_c, dx, kappa_hard):
    new = []
    for px, py in cp:
        dirs = set()
        sw = sx = sy = 0.0
        for (qx, qy), v, k in zip(pos_c, val_c, dir_c):
            if d32(qx, qy, px, py) <= F32(dx):
                dirs.add(int(k))
                w = math.exp(float(v))
                sw += w
                sx += w * float(qx)
                sy += w * float(qy)
        if len(dirs) <= kappa_hard:
            new.append((float(px), float(py)))
        else:
            new.append((sx / sw, sy / sw))
    return np.array(new)


def resample_loops(pos_c, val_c, dir_c, cp, dx, f_rest, eps=1e-6):
    P = len(cp)
    cp_f = np.zeros((P, 7))
    for (qx, qy), v, k in zip(pos_c, val_c, dir_c):
        near = []
        for j, (px, py) in enumerate(cp):
            d = d32(px, py, qx, qy)
            if d <= F32(dx):
                near.append((j, dx - math.hypot(float(px) - float(qx), float(py) - float(qy))))
        sw = sum(w for _, w in near)
        if sw <= 0:
            continue
        for j, w in near:
            cp_f[j, int(k)] += w / sw * float(v)
    out = cp_f.copy()
    for r, (px, py) in enumerate(cp):
        for i in range(7):
            if cp_f[r, i] < eps:
                sw = 0.0
                for j, (qx, qy) in enumerate(cp):
                    if d32(qx, qy, px, py) <= F32(dx) and cp_f[j, i] < eps:
                        sw += dx - math.hypot(float(qx) - float(px), float(qy) - float(py))
                if sw > 0:
                    out[r, i] = dx / sw * float(f_rest[i])
    return cp_f, out


# ---------------------------------------------------------------- options: periodic box, r_res, two-level
# Written as searches over explicit alternatives (the three periodic images of each coordinate
# difference, the inner / outer test per point) rather than the oracle's select formulas.


def _images32(a, b, L):
    """fp32 a - b and its two neighbouring periodic images; the one of least magnitude."""
    d = F32(F32(a) - F32(b))
    if L is None:
        return d
    L32 = F32(L)
    cands = [d, F32(d - L32), F32(d + L32)]
    return min(cands, key=lambda v: abs(float(v)))


def d32p(ax, ay, bx, by, L=None):
    """fp32 distance with the nearest periodic image in each coordinate (L None: open)."""
    ddx = _images32(ax, bx, L)
    ddy = _images32(ay, by, L)
    return F32(np.sqrt(F32(F32(ddx * ddx) + F32(ddy * ddy))))


def _disp64(a, b, L):
    """FP64 displacement a - b to the nearest periodic image (L None: plain)."""
    d = float(a) - float(b)
    if L is None:
        return d
    L64 = float(F32(L))
    return min((d, d - L64, d + L64), key=abs)


def thin_dict_options(pos_c, cs, n_period=None, two_level=None):
    """Thinning keys by explicit cases: fine cell (fx, fy) = floor(fp32(x / cs)) (mod the cells per
    period in a periodic box); with two_level = (cx, cy, r_in) a point whose fp32 squared distance
    to the centre is below fp32(r_in^2) keys the fine cell as (fx // 2, fy // 2, 2 (fx % 2) + fy % 2),
    any other point the coarse cell as (fx // 2, fy // 2, 4).  Control points in ascending key
    order, members in component order, FP64 centroid."""
    cells = {}
    for c, (x, y) in enumerate(pos_c):
        fx = math.floor(F32(F32(x) / F32(cs)))
        fy = math.floor(F32(F32(y) / F32(cs)))
        if n_period is not None:
            fx %= n_period
            fy %= n_period
        if two_level is None:
            key = (fx, fy, 0)
        else:
            cx, cy, rin = two_level
            ex = F32(F32(x) - F32(cx))
            ey = F32(F32(y) - F32(cy))
            inner = F32(F32(ex * ex) + F32(ey * ey)) < F32(F32(rin) * F32(rin))
            key = (fx // 2, fy // 2, 2 * (fx % 2) + (fy % 2)) if inner else (fx // 2, fy // 2, 4)
        cells.setdefault(key, []).append(c)
    out = []
    for key in sorted(cells):
        members = cells[key]
        sx = sum(float(pos_c[c][0]) for c in members) / len(members)
        sy = sum(float(pos_c[c][1]) for c in members) / len(members)
        out.append((key, members, (sx, sy)))
    return out


def refine_loops_periodic(cp, pos_c, val_c, dir_c, dx, kappa_hard, L):
    """Refine in a periodic box: the components within dx of the nearest image; a moved point goes
    to p + sum w (q - p) / sum w with q - p the nearest-image displacement, wrapped into [0, L)."""
    L64 = float(F32(L))
    new = []
    for px, py in cp:
        dirs = set()
        sw = sx = sy = 0.0
        for (qx, qy), v, k in zip(pos_c, val_c, dir_c):
            if d32p(qx, qy, px, py, L) <= F32(dx):
                dirs.add(int(k))
                w = math.exp(float(v))
                sw += w
                sx += w * _disp64(qx, px, L)
                sy += w * _disp64(qy, py, L)
        if len(dirs) <= kappa_hard:
            new.append((float(px), float(py)))
        else:
            new.append(((float(px) + sx / sw) % L64, (float(py) + sy / sw) % L64))
    return np.array(new)


def resample_loops_options(pos_c, val_c, dir_c, cp, r, f_rest, L=None, eps=1e-6):
    """Scatter and gap-fill with resample radius r (weights r - d, gap-fill (r / sw) f_rest) and
    nearest-image distances in a periodic box of side L."""
    P = len(cp)
    cp_f = np.zeros((P, 7))
    for (qx, qy), v, k in zip(pos_c, val_c, dir_c):
        near = []
        for j, (px, py) in enumerate(cp):
            if d32p(px, py, qx, qy, L) <= F32(r):
                near.append((j, r - math.hypot(_disp64(px, qx, L), _disp64(py, qy, L))))
        sw = sum(w for _, w in near)
        if sw <= 0:
            continue
        for j, w in near:
            cp_f[j, int(k)] += w / sw * float(v)
    out = cp_f.copy()
    for a, (px, py) in enumerate(cp):
        for i in range(7):
            if cp_f[a, i] < eps:
                sw = 0.0
                for j, (qx, qy) in enumerate(cp):
                    if d32p(qx, qy, px, py, L) <= F32(r) and cp_f[j, i] < eps:
                        sw += r - math.hypot(_disp64(qx, px, L), _disp64(qy, py, L))
                if sw > 0:
                    out[a, i] = r / sw * float(f_rest[i])
    return cp_f, out


def spawn_loops(cp, pos_c, dir_c, dx, kappa_hard, kappa_soft, C, offset):
    """Soft_outer spawn by explicit loops (direbm/reference/simulator.py:141-166): kappa and the
    vector sum of the FP64 directions C[k] over the components within dx of each control point;
    a point with kappa_hard < kappa <= kappa_soft and a nonzero sum spawns at cp + unit(sum) *
    offset * dx.  Returns [(parent index, (x, y))]."""
    out = []
    for r, (px, py) in enumerate(cp):
        dirs = set()
        sx = sy = 0.0
        for (qx, qy), k in zip(pos_c, dir_c):
            if d32p(qx, qy, px, py) <= F32(dx):
                dirs.add(int(k))
                sx += float(C[int(k)][0])
                sy += float(C[int(k)][1])
        n = math.hypot(sx, sy)
        if kappa_hard < len(dirs) <= kappa_soft and n > 1e-8:
            out.append((r, (float(px) + sx / n * offset * dx, float(py) + sy / n * offset * dx)))
    return out
