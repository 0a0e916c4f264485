"""Device plumbing: PyTorch-ROCm supplies device memory and the current HIP stream."""

from __future__ import annotations

import numpy as np

from ._lib import NoDeviceError

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t

        _torch = t
    return _torch


def default_device() -> str:
    """'cuda:0' when a HIP GPU is visible, else 'cpu' (direbm/warp/physics.py:21-22).

    The engine has no CPU path: compute calls on 'cpu' raise NoDeviceError.
    """
    t = torch()
    return "cuda:0" if t.cuda.is_available() else "cpu"


def require_gpu(device=None):
    t = torch()
    dev = device or default_device()
    if not str(dev).startswith("cuda") or not t.cuda.is_available():
        raise NoDeviceError("direbm_amd needs a HIP GPU (there is no CPU fallback)")
    return t.device(dev)


def stream_ptr(device) -> int:
    return int(torch().cuda.current_stream(device).cuda_stream)


def to_device(a, dtype, device):
    """numpy / DeviceArray / torch -> contiguous torch tensor on `device`."""
    t = torch()
    if isinstance(a, DeviceArray):
        a = a.tensor
    if isinstance(a, t.Tensor):
        return a.to(device=device, dtype=dtype).contiguous()
    return t.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(device).contiguous()


class DeviceArray:
    """Device array handed back by the stage functions: ``.shape``, ``.numpy()``, ``.tensor``.

    Stands in for the wp.array values the reference returns (direbm/warp/propagation.py:115-275).
    """

    def __init__(self, tensor):
        self.tensor = tensor

    @property
    def shape(self):
        return tuple(self.tensor.shape)

    def numpy(self):
        return self.tensor.detach().cpu().numpy()

    def __len__(self):
        return self.tensor.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.tensor.dtype}, device={self.tensor.device})"
