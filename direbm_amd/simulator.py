"""GpuSimulator: the DiReBM step on one MI355X (drop-in for direbm/warp/simulator.py:25-77).

State lives in engine-owned device memory (SoA, double-buffered, grow-only); ``step()`` runs
collide -> disperse -> thin -> refine -> resample with one host synchronisation per step (the
reference also synchronises once per step, at propagation.py:146).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _host
from ._device import require_gpu, stream_ptr
from ._lib import Params, call


def make_params(tau, dx=_host.DX, alpha=_host.ALPHA, kappa_hard=_host.KAPPA_HARD, rho_rest=1.0, r_res=None,
                box=None, two_level=None, eps=None, circle=None, kappa_soft=None) -> Params:
    """drb_params from GpuSimulator keyword arguments plus the engine's extensions.

    r_res: resample radius (default dx).  box: periodic square [0, box)^2.  two_level:
    (cx, cy, r_in) -- cells of 2*cs outside radius r_in of (cx, cy).  kappa_soft: the soft_outer
    spawn for kappa_hard < kappa <= kappa_soft (None: off, the reference GPU step).  circle is not a
    parameter of the engine's create call: it is set afterwards (set_obstacle).
    """
    p = Params()
    p.tau = float(tau)
    p.dx = float(dx)
    p.alpha = float(alpha)
    p.kappa_hard = int(kappa_hard)
    p.rho_rest = float(rho_rest)
    p.r_res = float(r_res) if r_res is not None else 0.0
    p.eps = float(eps) if eps is not None else 0.0
    if box is not None:
        p.periodic = 1
        p.box = float(box)
    if two_level is not None:
        p.two_level = 1
        p.cx, p.cy, p.r_in = (float(v) for v in two_level)
    p.kappa_soft = int(kappa_soft) if kappa_soft is not None else 0
    return p


def set_obstacle(h, inst, circle):
    """Circular obstacle (cx, cy, R) of instance `inst` of an engine handle."""
    cx, cy, r = (float(v) for v in circle)
    if not r > 0.0:
        raise ValueError("circle radius must be > 0")
    call("drb_sim_set_obstacle", h, int(inst), cx, cy, r)


class _StateView:
    """``sim.pos_m`` / ``sim.f_m``: ``.shape`` and ``.numpy()`` like the reference's wp.arrays."""

    def __init__(self, sim, which):
        self._sim, self._which = sim, which

    @property
    def shape(self):
        m = self._sim.count()
        return (m, 2) if self._which == 0 else (m, 7)

    def numpy(self):
        return self._sim.moments()[self._which]


class GpuSimulator:
    """DiReBM simulation of one point cloud (direbm/warp/simulator.py:25-48).

    GpuSimulator(pos_m (M,2), f_m (M,7), tau, *, dx=1, alpha=4, kappa_hard=4, rho_rest=1,
    device=None, r_res=None, box=None, two_level=None, circle=None, kappa_soft=None)

    kappa_soft: the per-object solver's soft_outer spawn (direbm/reference/simulator.py:141-166,
    KAPPA_SOFT = 5 in direbm/constants.py): a control point with kappa_hard < kappa <= kappa_soft
    adds a point 2 (1 - sqrt(3)/2) dx along the summed directions of its components.  None (the
    default) is the reference GPU step, which treats such points as inner.  Not with circle, graph
    mode or the adjoint's recorded steps.

    circle=(cx, cy, R): a circular obstacle, fluid outside (the reference's Circle,
    direbm/reference/boundary.py:20-47, with its specular bounce, simulator.py:90-117, and the
    removal of moments inside after each step, :79-80).  Not supported in graph mode or in the
    recorded steps of the adjoint.
    """

    def __init__(self, pos_m, f_m, tau, *, dx=_host.DX, alpha=_host.ALPHA, kappa_hard=_host.KAPPA_HARD,
                 rho_rest=1.0, device=None, r_res=None, box=None, two_level=None, eps=None, circle=None,
                 kappa_soft=None):
        dev = require_gpu(device)
        self.device = str(dev)
        self._dev = dev
        self.tau = float(tau)
        self.dx = float(dx)
        self.cell = float(dx / alpha)
        self.kappa_hard = int(kappa_hard)
        self.rho_rest = float(rho_rest)
        self.params = make_params(tau, dx, alpha, kappa_hard, rho_rest, r_res, box, two_level, eps,
                                  kappa_soft=kappa_soft)
        self._kw = dict(tau=float(tau), dx=float(dx), alpha=float(alpha), kappa_hard=int(kappa_hard),
                        rho_rest=float(rho_rest), r_res=r_res, box=box,
                        two_level=None if two_level is None else [float(v) for v in two_level], eps=eps,
                        circle=None if circle is None else [float(v) for v in circle],
                        kappa_soft=None if kappa_soft is None else int(kappa_soft))
        self._h = C.c_void_p()
        call("drb_sim_create", C.byref(self._h), C.byref(self.params), 1, stream_ptr(dev))
        if circle is not None:
            set_obstacle(self._h, 0, circle)
        self.iteration = 0
        self._tiles = (4, 4, 4, 4)  # refine lanes, depth, gather lanes, depth (the engine's default)
        self.set_state(pos_m, f_m)

    # --------------------------------------------------------------------------- state
    def set_state(self, pos_m, f_m):
        pos = np.ascontiguousarray(np.asarray(pos_m, np.float32).reshape(-1, 2))
        f = np.ascontiguousarray(np.asarray(f_m, np.float32).reshape(-1, 7))
        if len(pos) != len(f):
            raise ValueError("pos_m and f_m must have the same number of rows")
        call("drb_sim_set_state", self._h, pos.ctypes.data, f.ctypes.data, len(pos), None)

    def count(self) -> int:
        m = C.c_int64(0)
        call("drb_sim_count", self._h, C.byref(m), None)
        return int(m.value)

    @property
    def pos_m(self):
        return _StateView(self, 0)

    @property
    def f_m(self):
        return _StateView(self, 1)

    def moments(self):
        """(positions (M, 2), f (M, 7)) as numpy float32 (direbm/warp/simulator.py:75-77)."""
        m = self.count()
        pos = np.empty((m, 2), np.float32)
        f = np.empty((m, 7), np.float32)
        if m:
            call("drb_sim_get_state", self._h, pos.ctypes.data, f.ctypes.data)
        return pos, f

    # ---------------------------------------------------------------------------- step
    def step(self, nsteps: int = 1):
        """Advance nsteps steps (direbm/warp/simulator.py:50-73)."""
        call("drb_sim_step", self._h, int(nsteps))
        self.iteration += int(nsteps)

    def save(self, path):
        """Write a checkpoint (.npz: positions, distributions, iteration, parameters)."""
        from .checkpoint import save

        pos, f = self.moments()
        t = self._tiles
        save(path, pos, f, self.iteration, [self._kw], [len(pos)], (t[0], t[2]), (t[1], t[3]))

    @classmethod
    def load(cls, path, device=None):
        """A GpuSimulator resumed from a checkpoint written by ``save``."""
        from .checkpoint import load

        pos, f, iteration, kws, _, lanes, depth = load(path)
        if len(kws) != 1:
            raise ValueError(f"{path} holds {len(kws)} instances; use BatchSimulator.load")
        sim = cls(pos, f, device=device, **kws[0])
        sim.set_tiles(lanes[0], depth[0], lanes[1], depth[1])
        sim.iteration = iteration
        return sim

    def diagnostics(self):
        """FP64 mass, momentum, and the last step's dropped / injected mass."""
        out = np.zeros(5, np.float64)
        call("drb_sim_diagnostics", self._h, out.ctypes.data)
        return {"mass": out[0], "mom_x": out[1], "mom_y": out[2], "dropped": out[3], "injected": out[4]}

    def set_timing(self, on=True, stages=None):
        """Time the stages of each step with HIP events (stage_ms); `stages` (names from STAGES)
        times only those, with fewer events in the step."""
        if on and stages is not None:
            mask = 0
            for name in stages:
                mask |= 1 << self.STAGES.index(name)
            call("drb_sim_set_timing", self._h, 0x200 | mask)
        else:
            call("drb_sim_set_timing", self._h, 1 if on else 0)

    def autotune(self):
        """Advance one step and choose the tiles (lanes per point, candidates in flight per lane)
        of the refine and gather kernels by timing five of them on that step (see
        drb_sim_autotune).  Returns the choice."""
        lanes = np.zeros(2, np.int32)
        call("drb_sim_autotune", self._h, lanes.ctypes.data)
        self.iteration += 1
        return self.tiles()

    def tiles(self):
        """The refine and gather tiles: {refine,gather}_{lanes,depth}."""
        t = np.zeros(4, np.int32)
        call("drb_sim_get_tiles", self._h, t.ctypes.data)
        self._tiles = tuple(int(v) for v in t)
        return {"refine_lanes": self._tiles[0], "refine_depth": self._tiles[1], "gather_lanes": self._tiles[2],
                "gather_depth": self._tiles[3]}

    def set_lanes(self, refine_lanes=4, gather_lanes=4):
        """Pin the refine / gather lanes per point (2, 4 or 8), e.g. to reproduce a tuned run."""
        call("drb_sim_set_lanes", self._h, int(refine_lanes), int(gather_lanes))
        self.tiles()

    def set_tiles(self, refine_lanes=4, refine_depth=4, gather_lanes=4, gather_depth=4):
        """Pin the refine / gather tiles: lanes per point (2, 4 or 8) and candidates in flight per
        lane (2 or 4)."""
        call("drb_sim_set_tiles", self._h, int(refine_lanes), int(refine_depth), int(gather_lanes), int(gather_depth))
        self.tiles()

    def set_graph(self, on=True, horizon=None):
        """Step through captured HIP graphs: counts kept on the device, one host synchronisation
        per ``step(n)`` call.  A periodic grid is fixed; an open domain's grid is laid out on the
        device each step within a cell capacity sized for ``horizon`` steps of the largest growth
        (default 16); a step beyond it resumes the call from a new capture.  Results are
        bit-identical to plain steps."""
        if horizon is not None:
            call("drb_sim_set_graph_horizon", self._h, int(horizon))
        call("drb_sim_set_graph", self._h, 1 if on else 0)

    def graph_info(self):
        """Graph captures so far, replays resumed after an empty step, and the capacity of the
        current capture (cells, control points)."""
        out = (C.c_int64 * 4)()
        call("drb_sim_graph_info", self._h, out)
        return {"captures": int(out[0]), "resumes": int(out[1]), "cells": int(out[2]), "points": int(out[3])}

    STAGES = ("collide", "keys", "sort", "thin", "refine", "bin", "norm", "gather", "gapfill")

    @property
    def updates(self) -> int:
        """Point-updates performed so far: the moments entering each step, summed over steps."""
        v = C.c_int64(0)
        call("drb_sim_updates", self._h, C.byref(v))
        return int(v.value)

    def stage_ms(self):
        """Per-stage GPU time (ms) per step, averaged over the steps of the last step() call, from
        HIP events on the engine's stream."""
        out = np.zeros(len(self.STAGES), np.float32)
        call("drb_sim_stage_ms", self._h, out.ctypes.data)
        return dict(zip(self.STAGES, out.tolist()))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            call("drb_sim_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["GpuSimulator", "make_params"]
