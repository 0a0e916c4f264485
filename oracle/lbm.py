"""FP64 D2Q7 hexagonal LBM (oracle).  Restates direbm/lbm.py:26-66.

Nodes on skewed axial coordinates (i, j) -> physical i*(1,0) + j*(1/2, sqrt(3)/2); direction k
streams by the integer shift HEX_OFFSETS[k]; periodic wrap; collide then stream each step.
"""

from __future__ import annotations

import numpy as np

from .constants import HEX_OFFSETS
from .physics import collide, equilibrium, recover

_S = np.sqrt(3.0) / 2.0


class HexLBM:
    def __init__(self, ni, nj, tau, rho0=1.0):
        self.ni, self.nj, self.tau = int(ni), int(nj), float(tau)
        self.f = equilibrium(np.full((self.nj, self.ni), float(rho0)), np.zeros((self.nj, self.ni, 2)))
        self.center = (self.ni // 2, self.nj // 2)

    def set_pulse(self, rho_pulse):
        """direbm/lbm.py:45-47: equilibrium at rest with density rho_pulse at the centre node."""
        ci, cj = self.center
        self.f[cj, ci] = equilibrium(np.float64(rho_pulse), np.zeros(2))

    def set_density(self, rho):
        """Rest equilibrium of a given (nj, ni) density field (used for the Gaussian-pulse case)."""
        rho = np.asarray(rho, np.float64)
        self.f = equilibrium(rho, np.zeros(rho.shape + (2,)))

    def stream(self):
        new = np.empty_like(self.f)
        for k, (di, dj) in enumerate(HEX_OFFSETS):
            new[..., k] = np.roll(self.f[..., k], shift=(dj, di), axis=(0, 1))
        self.f = new

    def step(self):
        self.f = collide(self.f, self.tau)
        self.stream()

    def macroscopic(self):
        return recover(self.f)

    def radius(self):
        ii, jj = np.meshgrid(np.arange(self.ni), np.arange(self.nj))
        ci, cj = self.center
        x = (ii + jj * 0.5) - (ci + cj * 0.5)
        y = (jj - cj) * _S
        return np.sqrt(x * x + y * y)
