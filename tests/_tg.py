"""Observables of the periodic Taylor-Green runs used by the long-horizon band test (shared by
the test and by the script that writes its fixture)."""

from __future__ import annotations

import numpy as np

from oracle.constants import C2

NBINS = 8  # bins per side: 16 h at 128 x 128 seeding cells


def observables(pos, f, L, nbins=NBINS):
    """Mode amplitude, FP64 mass and binned density / velocity of a Taylor-Green state.

    Amplitude: the mass-weighted projection of the points' velocities onto the initial mode
    phi = (-cos kx sin ky, sin kx cos ky), k = 2 pi / L: sum m_i u_i . phi_i / sum m_i |phi_i|^2
    (u0 for the initial field).  Binned fields on nbins x nbins bins of [0, L)^2."""
    pos = np.asarray(pos, np.float64)
    f = np.asarray(f, np.float64)
    m = f.sum(1)
    mom = f @ C2
    k = 2 * np.pi / L
    x, y = k * pos[:, 0], k * pos[:, 1]
    phx, phy = -np.cos(x) * np.sin(y), np.sin(x) * np.cos(y)
    amp = float((mom[:, 0] * phx + mom[:, 1] * phy).sum() / (m * (phx * phx + phy * phy)).sum())
    edges = np.linspace(0.0, L, nbins + 1)
    bm, _, _ = np.histogram2d(pos[:, 0], pos[:, 1], bins=[edges, edges], weights=m)
    bx, _, _ = np.histogram2d(pos[:, 0], pos[:, 1], bins=[edges, edges], weights=mom[:, 0])
    by, _, _ = np.histogram2d(pos[:, 0], pos[:, 1], bins=[edges, edges], weights=mom[:, 1])
    area = (L / nbins) ** 2
    return {"amp": amp, "mass": float(m.sum()), "rho": bm / area, "ux": bx / bm, "uy": by / bm}
