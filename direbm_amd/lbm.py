"""GpuHexLBM: D2Q7 hexagonal LBM on the GPU (drop-in for direbm/warp/lbm.py:72-115).

Populations are SoA planes f[k][j][i] in HBM (ping-pong pair); one fused collide+push-stream
kernel per step.  ``.f.numpy()`` returns the reference's (nj, ni, 7) layout.
"""

from __future__ import annotations

import numpy as np

from . import _host
from ._device import require_gpu, stream_ptr, torch
from ._lib import call

_S = np.sqrt(3.0) / 2.0


class _Populations:
    def __init__(self, lbm):
        self._lbm = lbm

    @property
    def shape(self):
        return (self._lbm.nj, self._lbm.ni, 7)

    def numpy(self):
        return self._lbm.f_soa.permute(1, 2, 0).contiguous().cpu().numpy()


class GpuHexLBM:
    def __init__(self, ni, nj, tau, rho0=1.0, device=None):
        dev = require_gpu(device)
        self.device = str(dev)
        self._dev = dev
        self.ni, self.nj, self.tau = int(ni), int(nj), float(tau)
        t = torch()
        # FP64 numpy equilibrium cast to fp32, as direbm/warp/lbm.py:78-79.
        f0 = _host.equilibrium(np.full((self.nj, self.ni), float(rho0)), np.zeros((self.nj, self.ni, 2)))
        self.f_soa = t.as_tensor(np.ascontiguousarray(f0.astype(np.float32).transpose(2, 0, 1))).to(dev)
        self.out_soa = t.zeros_like(self.f_soa)
        self.center = (self.ni // 2, self.nj // 2)

    @property
    def f(self):
        return _Populations(self)

    def set_density(self, rho):
        """Rest equilibrium of an (nj, ni) density field (FP64 on the host, cast to fp32)."""
        rho = np.asarray(rho, np.float64).reshape(self.nj, self.ni)
        f0 = _host.equilibrium(rho, np.zeros((self.nj, self.ni, 2)))
        self.f_soa.copy_(torch().as_tensor(np.ascontiguousarray(f0.astype(np.float32).transpose(2, 0, 1))))

    def set_pulse(self, rho_pulse):
        """f[cj, ci, k] = rho W_k (direbm/warp/lbm.py:86-93)."""
        ci, cj = self.center
        call("drb_lbm_set_node", self.f_soa.data_ptr(), self.ni, self.nj, ci, cj, float(rho_pulse),
             stream_ptr(self._dev))

    def step(self, nsteps: int = 1):
        """Collide then periodic stream (direbm/warp/lbm.py:95-105), nsteps times."""
        n = int(nsteps)
        call("drb_lbm_step", self.f_soa.data_ptr(), self.out_soa.data_ptr(), self.ni, self.nj, self.tau, n,
             stream_ptr(self._dev))
        if n % 2:
            self.f_soa, self.out_soa = self.out_soa, self.f_soa

    def macroscopic(self):
        """(rho (nj, ni), u (nj, ni, 2)) in FP64 on the host (direbm/warp/lbm.py:107-108)."""
        return _host.recover(self.f.numpy())

    def radius(self):
        ii, jj = np.meshgrid(np.arange(self.ni), np.arange(self.nj))
        ci, cj = self.center
        x = (ii + jj * 0.5) - (ci + cj * 0.5)
        y = (jj - cj) * _S
        return np.sqrt(x * x + y * y)


__all__ = ["GpuHexLBM"]
