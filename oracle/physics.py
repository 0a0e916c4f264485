"""FP64 BGK physics (oracle).  Restates direbm/physics.py:16-55.

Arrays are shaped (..., Q) for distributions.  ``recover`` guards rho == 0 (u = 0), as
direbm/physics.py:24 does; the HIP kernels follow the same guard.
"""

from __future__ import annotations

import numpy as np

from .constants import C2, CS2_2D, W2


def recover(f, C=C2):
    """rho = sum_i f_i, u = sum_i f_i c_i / rho (u = 0 where rho == 0).  direbm/physics.py:16-26."""
    f = np.asarray(f, dtype=np.float64)
    rho = f.sum(axis=-1)
    mom = f @ C
    safe = np.where(rho == 0.0, 1.0, rho)
    return rho, mom / safe[..., None]


def equilibrium(rho, u, C=C2, W=W2, cs2=CS2_2D):
    """f_eq_i = rho W_i (1 + cu/cs2 + cu^2/(2 cs2^2) - u.u/(2 cs2)).  direbm/physics.py:29-45."""
    rho = np.asarray(rho, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    inv = 1.0 / cs2
    cu = u @ C.T
    usq = (u * u).sum(axis=-1)
    return rho[..., None] * W * (1.0 + inv * cu + 0.5 * inv * inv * cu * cu - 0.5 * inv * usq[..., None])


def collide(f, tau, C=C2, W=W2, cs2=CS2_2D):
    """f* = f + (f_eq(rho, u) - f) / tau.  direbm/physics.py:48-55."""
    f = np.asarray(f, dtype=np.float64)
    rho, u = recover(f, C)
    return f + (equilibrium(rho, u, C, W, cs2) - f) / tau
