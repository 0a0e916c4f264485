"""Dispersion, cell-thinning, refine and resample stages on the GPU.

Drop-in for the stage functions of direbm/warp/propagation.py:115-275: same names, same argument
meaning, device arrays back (``DeviceArray`` with ``.shape`` / ``.numpy()``).  Each call runs the
HIP kernels of libdirebm_hip.so through the C-ABI; the neighbour structure is a sorted cell grid
built per call (the reference builds a wp.HashGrid per call).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ._device import DeviceArray, require_gpu, stream_ptr, to_device, torch
from ._lib import call

_STAGES = {}


def _stage(dev):
    key = str(dev)
    if key not in _STAGES:
        h = C.c_void_p()
        call("drb_stage_create", C.byref(h))
        _STAGES[key] = h
    return _STAGES[key]


def _f32(a, dev, cols=None):
    t = to_device(a, torch().float32, dev)
    return t.reshape(-1, cols) if cols else t.reshape(-1)


def disperse(pos_m, f_m, dx, device=None):
    """Moments -> components (pos_c (7M, 2), val_c (7M,), dir_c (7M,) int32).

    direbm/warp/propagation.py:115-127: pos_c[7i+k] = pos_m[i] + C[k]*dx, val_c = f_m[i,k].
    """
    dev = require_gpu(device)
    t = torch()
    pm = _f32(pos_m, dev, 2)
    fm = _f32(f_m, dev, 7)
    m = pm.shape[0]
    pos_c = t.empty((7 * m, 2), dtype=t.float32, device=dev)
    val_c = t.empty((7 * m,), dtype=t.float32, device=dev)
    dir_c = t.empty((7 * m,), dtype=t.int32, device=dev)
    call("drb_disperse", pm.data_ptr(), fm.data_ptr(), m, float(dx), pos_c.data_ptr(), val_c.data_ptr(),
         dir_c.data_ptr(), stream_ptr(dev))
    return DeviceArray(pos_c), DeviceArray(val_c), DeviceArray(dir_c)


def create_control_points(pos_c, cs, device=None):
    """Cell-thin a component cloud -> (cp_pos (P, 2), P).  direbm/warp/propagation.py:130-153."""
    dev = require_gpu(device)
    t = torch()
    pc = _f32(pos_c, dev, 2)
    nc = pc.shape[0]
    cp = t.empty((max(nc, 1), 2), dtype=t.float32, device=dev)
    p = C.c_int64(0)
    call("drb_create_control_points", _stage(dev), pc.data_ptr(), nc, float(cs), cp.data_ptr(), nc, C.byref(p),
         stream_ptr(dev))
    return DeviceArray(cp[: p.value]), int(p.value)


def refine_control_points(cp_pos, pos_c, val_c, dir_c, dx, kappa_hard, device=None):
    """Refined control-point positions (P, 2).  direbm/warp/propagation.py:249-275."""
    dev = require_gpu(device)
    t = torch()
    cp = _f32(cp_pos, dev, 2)
    pc = _f32(pos_c, dev, 2)
    vc = _f32(val_c, dev)
    dc = to_device(dir_c, t.int32, dev).reshape(-1)
    out = t.empty_like(cp)
    call("drb_refine_control_points", _stage(dev), cp.data_ptr(), cp.shape[0], pc.data_ptr(), vc.data_ptr(),
         dc.data_ptr(), pc.shape[0], float(dx), int(kappa_hard), out.data_ptr(), stream_ptr(dev))
    return DeviceArray(out)


def resample(pos_c, val_c, dir_c, cp_pos, dx, f_rest, device=None):
    """Scatter + gap-fill/emit -> (out_pos (P, 2), out_f (P, 7)).  direbm/warp/propagation.py:221-246."""
    dev = require_gpu(device)
    t = torch()
    cp = _f32(cp_pos, dev, 2)
    pc = _f32(pos_c, dev, 2)
    vc = _f32(val_c, dev)
    dc = to_device(dir_c, t.int32, dev).reshape(-1)
    fr = _f32(np.asarray(f_rest, np.float32) if not hasattr(f_rest, "tensor") else f_rest, dev)
    p = cp.shape[0]
    out_pos = t.empty((p, 2), dtype=t.float32, device=dev)
    out_f = t.empty((p, 7), dtype=t.float32, device=dev)
    call("drb_resample", _stage(dev), pc.data_ptr(), vc.data_ptr(), dc.data_ptr(), pc.shape[0], cp.data_ptr(), p,
         float(dx), fr.data_ptr(), out_pos.data_ptr(), out_f.data_ptr(), stream_ptr(dev))
    return DeviceArray(out_pos), DeviceArray(out_f)


__all__ = ["disperse", "create_control_points", "refine_control_points", "resample"]
