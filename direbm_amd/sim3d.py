"""GpuSimulator3D: the DiReBM step in 3D on the D3Q13 icosahedral lattice, on one GPU.

No reference GPU counterpart: the reference runs 3D only in its per-object CPU solver
(direbm/reference/simulator.py:30-198 with lattice=D3Q13) and plans the GPU step for 3D as the 2D
step with vec3 positions and 13 directions (HANDOFF.md:135-172).  The interface mirrors
GpuSimulator (direbm/warp/simulator.py:25-77) with (M, 3) positions and (M, 13) distributions.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ._device import require_gpu, stream_ptr
from ._lib import call
from .simulator import make_params


def d3q13():
    """The fp32 D3Q13 velocity set (13, 3) and weights (13,) the 3D kernels use."""
    c = np.zeros(39, np.float32)
    w = np.zeros(13, np.float32)
    call("drb_d3q13", c.ctypes.data, w.ctypes.data)
    return c.reshape(13, 3), w


class GpuSimulator3D:
    """GpuSimulator3D(pos_m (M,3), f_m (M,13), tau, *, dx=1, alpha=4, kappa_hard=4, rho_rest=1,
    r_res=None, box=None, sphere=None, eps=None, device=None, kappa_soft=None).

    kappa_soft is refused (DrbError): the reference's soft_outer spawn is 2D only
    (direbm/reference/simulator.py:7-9); the argument exists so that 2D and 3D keywords match.

    sphere: (cx, cy, cz, R), a spherical obstacle with the fluid outside
    (direbm/reference/boundary.py:50-51 with direbm/reference/simulator.py:79-117): components
    that would land inside bounce specularly and split over the lattice directions, and points
    inside it after resampling are dropped.  In a periodic box it must lie 2 dx inside the box."""

    def __init__(self, pos_m, f_m, tau, *, dx=1.0, alpha=4.0, kappa_hard=4, rho_rest=1.0, r_res=None, box=None,
                 sphere=None, eps=None, device=None, kappa_soft=None):
        dev = require_gpu(device)
        self.device = str(dev)
        self.tau, self.dx, self.cell = float(tau), float(dx), float(dx / alpha)
        self.params = make_params(tau, dx, alpha, kappa_hard, rho_rest, r_res, box, None, eps, kappa_soft=kappa_soft)
        self._h = C.c_void_p()
        call("drb_sim3_create", C.byref(self._h), C.byref(self.params), stream_ptr(dev))
        self.sphere = None if sphere is None else tuple(float(v) for v in sphere)
        if self.sphere is not None:
            call("drb_sim3_set_sphere", self._h, *self.sphere)
        self.iteration = 0
        self.set_state(pos_m, f_m)

    def set_state(self, pos_m, f_m):
        pos = np.ascontiguousarray(np.asarray(pos_m, np.float32).reshape(-1, 3))
        f = np.ascontiguousarray(np.asarray(f_m, np.float32).reshape(-1, 13))
        if len(pos) != len(f):
            raise ValueError("pos_m and f_m must have the same number of rows")
        call("drb_sim3_set_state", self._h, pos.ctypes.data, f.ctypes.data, len(pos))

    def set_lanes(self, refine_lanes=4, gather_lanes=4):
        """Lanes per control point of the refine and gather kernels (4, 8 or 16)."""
        call("drb_sim3_set_lanes", self._h, int(refine_lanes), int(gather_lanes))

    def count(self) -> int:
        m = C.c_int64(0)
        call("drb_sim3_count", self._h, C.byref(m))
        return int(m.value)

    def moments(self):
        m = self.count()
        pos = np.empty((m, 3), np.float32)
        f = np.empty((m, 13), np.float32)
        if m:
            call("drb_sim3_get_state", self._h, pos.ctypes.data, f.ctypes.data)
        return pos, f

    def step(self, nsteps: int = 1):
        call("drb_sim3_step", self._h, int(nsteps))
        self.iteration += int(nsteps)

    def diagnostics(self):
        out = np.zeros(4, np.float64)
        call("drb_sim3_diagnostics", self._h, out.ctypes.data)
        return {"mass": out[0], "mom_x": out[1], "mom_y": out[2], "mom_z": out[3]}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            call("drb_sim3_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["GpuSimulator3D", "d3q13"]
