"""Host-side FP64 helpers the reference itself runs in numpy around its GPU kernels.

The reference initialises GpuHexLBM with numpy's FP64 equilibrium and reads macroscopic fields
back with numpy's recover (direbm/warp/lbm.py:14-15,78,107-108), and builds f_rest = rho_rest * W
on the host (direbm/warp/simulator.py:47).  These helpers do the same host-side conversions; no
step of the simulation runs here.
"""

from __future__ import annotations

import numpy as np

_S = np.sqrt(3.0) / 2.0
C = np.array([[0.0, 0.0], [1.0, 0.0], [0.5, _S], [-0.5, _S], [-1.0, 0.0], [-0.5, -_S], [0.5, -_S]])
W = np.array([0.5] + [1.0 / 12.0] * 6)
CS2 = 0.25
DX = 1.0
ALPHA = 4.0
KAPPA_HARD = 4


def equilibrium(rho, u):
    """FP64 f_eq (direbm/physics.py:29-45), shape (..., 7)."""
    rho = np.asarray(rho, np.float64)
    u = np.asarray(u, np.float64)
    inv = 1.0 / CS2
    cu = u @ C.T
    usq = (u * u).sum(axis=-1)
    return rho[..., None] * W * (1.0 + inv * cu + 0.5 * inv * inv * cu * cu - 0.5 * inv * usq[..., None])


def recover(f):
    """FP64 (rho, u) with the rho == 0 guard (direbm/physics.py:16-26)."""
    f = np.asarray(f, np.float64)
    rho = f.sum(axis=-1)
    mom = f @ C
    safe = np.where(rho == 0.0, 1.0, rho)
    return rho, mom / safe[..., None]
