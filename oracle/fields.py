"""Macroscopic fields from scattered moments (oracle).  Restates direbm/fields.py:17-41.

Mass and momentum of each moment are histogrammed onto a square grid of cell size h over
[xmin, xmax]^2; density = cell mass / h^2, velocity = cell momentum / cell mass (0 where empty).
Arrays are indexed [iy, ix].
"""

from __future__ import annotations

import numpy as np

from .constants import C2


def bin_fields(x, f, xmin, xmax, h):
    x = np.asarray(x, dtype=np.float64)
    f = np.asarray(f, dtype=np.float64)
    edges = np.arange(xmin, xmax + h, h)
    mass = f.sum(axis=1)
    mom = f @ C2
    m, _, _ = np.histogram2d(x[:, 0], x[:, 1], bins=[edges, edges], weights=mass)
    px, _, _ = np.histogram2d(x[:, 0], x[:, 1], bins=[edges, edges], weights=mom[:, 0])
    py, _, _ = np.histogram2d(x[:, 0], x[:, 1], bins=[edges, edges], weights=mom[:, 1])
    m, px, py = m.T, px.T, py.T
    rho = m / (h * h)
    safe = np.where(m == 0.0, 1.0, m)
    u = np.stack([px / safe, py / safe], axis=-1)
    u[m == 0.0] = 0.0
    return rho, u, (xmin, xmax, xmin, xmax)
