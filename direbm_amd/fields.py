"""Macroscopic fields from scattered moments on the GPU (drop-in for direbm/fields.py:17-41).

``bin_fields`` bins each moment's mass sum_k f_k and momentum sum_k f_k c_k onto a square grid
(FP64 sums in ascending moment order, the order numpy's histogram2d adds its weights in) and
returns (rho, u, extent) as the reference does.  ``bin_moments`` is the differentiable torch form
(w.r.t. f) used by the adjoint: it returns the raw FP64 mass and momentum grids.
"""

from __future__ import annotations

import numpy as np

from ._device import require_gpu, stream_ptr, to_device, torch
from ._lib import call
from .propagation import _stage


def bin_edges(xmin, xmax, h):
    """The reference's bin edges np.arange(xmin, xmax + h, h) (direbm/fields.py:27)."""
    return np.arange(xmin, xmax + h, h)


def _edges_t(edges, dev):
    return torch().as_tensor(np.ascontiguousarray(edges, np.float64), device=dev)


class _BinMoments:
    """torch.autograd.Function (built lazily so importing needs no torch)."""

    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            t = torch()

            class BinMoments(t.autograd.Function):
                @staticmethod
                def forward(ctx, pos, f, edges_t):
                    dev = pos.device
                    nb = edges_t.shape[0] - 1
                    mass = t.empty((nb, nb), dtype=t.float64, device=dev)
                    mom = t.empty((nb, nb, 2), dtype=t.float64, device=dev)
                    call("drb_bin_fields", _stage(dev), pos.data_ptr(), f.data_ptr(), pos.shape[0],
                         edges_t.data_ptr(), nb, mass.data_ptr(), mom.data_ptr(), stream_ptr(dev))
                    ctx.save_for_backward(pos, edges_t)
                    return mass, mom

                @staticmethod
                def backward(ctx, gmass, gmom):
                    pos, edges_t = ctx.saved_tensors
                    dev = pos.device
                    nb = edges_t.shape[0] - 1
                    gm = gmass.contiguous() if gmass is not None else None
                    gp = gmom.contiguous() if gmom is not None else None
                    gf = t.empty((pos.shape[0], 7), dtype=t.float32, device=dev)
                    call("drb_bin_fields_vjp", pos.data_ptr(), pos.shape[0], edges_t.data_ptr(), nb,
                         gm.data_ptr() if gm is not None else None, gp.data_ptr() if gp is not None else None,
                         gf.data_ptr(), stream_ptr(dev))
                    return None, gf, None

            cls.fn = BinMoments
        return cls.fn


def bin_moments(pos, f, edges):
    """(mass (nb, nb), momentum (nb, nb, 2)) FP64 torch tensors, [iy, ix]; differentiable w.r.t. f.

    pos (M, 2), f (M, 7): float32 torch tensors on the GPU; edges: 1-D bin edges (numpy).
    """
    dev = require_gpu(pos.device)
    t = torch()
    pos = pos.detach().to(t.float32).contiguous()
    f = f.to(t.float32).contiguous()
    return _BinMoments.get().apply(pos, f, _edges_t(edges, dev))


def bin_fields(x, f, xmin, xmax, h, device=None):
    """(rho (ny, nx), u (ny, nx, 2), extent) as numpy FP64 (direbm/fields.py:17-41)."""
    dev = require_gpu(device)
    t = torch()
    pos = to_device(np.asarray(x, np.float32).reshape(-1, 2) if not isinstance(x, t.Tensor) else x, t.float32, dev)
    ff = to_device(np.asarray(f, np.float32).reshape(-1, 7) if not isinstance(f, t.Tensor) else f, t.float32, dev)
    with t.no_grad():
        mass, mom = bin_moments(pos.reshape(-1, 2), ff.reshape(-1, 7), bin_edges(xmin, xmax, h))
    m = mass.cpu().numpy()
    p = mom.cpu().numpy()
    rho = m / (h * h)
    safe = np.where(m == 0.0, 1.0, m)
    u = p / safe[..., None]
    u[m == 0.0] = 0.0
    return rho, u, (xmin, xmax, xmin, xmax)


def fields_t(mass, mom, h):
    """Differentiable (rho, u) from bin_moments output: rho = mass / h^2, u = mom / mass (0 where
    empty), as direbm/fields.py:37-41."""
    t = torch()
    empty = mass == 0.0
    safe = t.where(empty, t.ones_like(mass), mass)
    u = mom / safe[..., None]
    u = t.where(empty[..., None], t.zeros_like(u), u)
    return mass / (h * h), u


__all__ = ["bin_fields", "bin_moments", "bin_edges", "fields_t"]
