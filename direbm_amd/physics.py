"""BGK collide and equilibrium on the GPU (drop-in for direbm/warp/physics.py:78-100)."""

from __future__ import annotations

import numpy as np

from ._device import default_device, require_gpu, stream_ptr, to_device, torch
from ._lib import call


def collide(f_np, tau, device=None):
    """BGK collision of f (N, 7) on the GPU; returns a new (N, 7) float32 numpy array.

    Same contract as direbm/warp/physics.py:78-85 (host in, host out, synchronised).
    """
    dev = require_gpu(device)
    f = to_device(np.asarray(f_np, np.float32).reshape(-1, 7), torch().float32, dev)
    call("drb_collide", f.data_ptr(), f.shape[0], 7, 1, float(tau), stream_ptr(dev))
    torch().cuda.synchronize(dev)
    return f.cpu().numpy()


def equilibrium(rho_np, u_np, device=None):
    """Equilibrium (N, 7) float32 of rho (N,), u (N, 2) (direbm/warp/physics.py:88-97)."""
    dev = require_gpu(device)
    t = torch()
    rho = to_device(np.asarray(rho_np, np.float32).reshape(-1), t.float32, dev)
    u = to_device(np.asarray(u_np, np.float32).reshape(-1, 2), t.float32, dev)
    out = t.empty((rho.shape[0], 7), dtype=t.float32, device=dev)
    call("drb_equilibrium", rho.data_ptr(), u.data_ptr(), rho.shape[0], out.data_ptr(), 7, 1, stream_ptr(dev))
    t.cuda.synchronize(dev)
    return out.cpu().numpy()


__all__ = ["collide", "equilibrium", "default_device"]
