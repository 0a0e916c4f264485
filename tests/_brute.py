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
