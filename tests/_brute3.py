"""Pure-Python loop restatements of the 3D step stages for tiny inputs: explicit loops over all
pairs (no k-d tree, no vectorisation), fp32 decisions written from the kernels' definition."""

from __future__ import annotations

import math

import numpy as np

F32 = np.float32


def d32(a, b):
    s = F32(0.0)
    dd = [F32(F32(a[k]) - F32(b[k])) for k in range(3)]
    s = F32(F32(dd[0] * dd[0]) + F32(dd[1] * dd[1]))
    s = F32(s + F32(dd[2] * dd[2]))
    return F32(np.sqrt(s))


def thin_dict(pos_c, cs):
    cells = {}
    for c, p in enumerate(pos_c):
        key = tuple(math.floor(F32(F32(p[k]) / F32(cs))) for k in range(3))
        cells.setdefault(key, []).append(c)
    out = []
    for key in sorted(cells):
        mem = cells[key]
        acc = [F32(0.0)] * 3
        for c in mem:
            acc = [F32(acc[k] + F32(pos_c[c][k])) for k in range(3)]
        out.append((key, mem, tuple(F32(acc[k] / F32(len(mem))) for k in range(3))))
    return out


def refine_loops(cp, pos_c, val_c, dir_c, dx, kappa_hard):
    out, kap = [], []
    for p in cp:
        mask, sw, s = 0, 0.0, [0.0, 0.0, 0.0]
        for q, v, d in zip(pos_c, val_c, dir_c):
            if d32(q, p) <= F32(dx):
                mask |= 1 << int(d)
                w = math.exp(float(v))
                sw += w
                s = [s[k] + w * float(q[k]) for k in range(3)]
        kappa = bin(mask).count("1")
        kap.append(kappa)
        out.append([s[k] / sw for k in range(3)] if kappa > kappa_hard else [float(p[k]) for k in range(3)])
    return np.array(out), np.array(kap)


def resample_loops(pos_c, val_c, dir_c, cp, r, f_rest, eps=1e-6):
    P = len(cp)
    cpf = np.zeros((P, 13))
    for q, v, d in zip(pos_c, val_c, dir_c):
        ws = []
        for j, p in enumerate(cp):
            if d32(p, q) <= F32(r):
                ws.append((j, r - math.dist([float(x) for x in p], [float(x) for x in q])))
        sw = sum(w for _, w in ws)
        if sw <= 0:
            continue
        for j, w in ws:
            cpf[j, int(d)] += w / sw * float(v)
    out = cpf.copy()
    for a in range(P):
        for i in range(13):
            if cpf[a, i] < eps:
                sw = 0.0
                for b in range(P):
                    if cpf[b, i] < eps and d32(cp[b], cp[a]) <= F32(r):
                        sw += r - math.dist([float(x) for x in cp[a]], [float(x) for x in cp[b]])
                out[a, i] = r / sw * f_rest[i]
    return cpf, out
