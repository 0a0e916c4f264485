"""BatchSimulator: many independent DiReBM simulations advanced together, one launch per stage.

The instances share one segment table on the device (per-instance parameters, cell grids and
moment ranges): keys carry each instance's cell offset, so one sort and one pass of every stage
kernel serve all instances, and each instance's result is bit-for-bit that of its own
GpuSimulator run.  No reference counterpart (the reference simulates one point cloud at a time).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ._device import require_gpu, stream_ptr
from ._lib import Params, call
from .simulator import make_params, set_obstacle


class BatchSimulator:
    """BatchSimulator(cases) with cases = [(pos (M_b,2), f (M_b,7), kwargs), ...]; kwargs as GpuSimulator."""

    def __init__(self, cases, device=None):
        dev = require_gpu(device)
        self.device = str(dev)
        self._dev = dev
        self.n = len(cases)
        if self.n < 1:
            raise ValueError("at least one instance")
        arr = (Params * self.n)()
        for b, (_, _, kw) in enumerate(cases):
            arr[b] = make_params(**kw)
        self._kws = [dict(kw) for _, _, kw in cases]
        self._params = arr
        self._h = C.c_void_p()
        call("drb_sim_create", C.byref(self._h), arr, self.n, stream_ptr(dev))
        for b, (_, _, kw) in enumerate(cases):
            if kw.get("circle") is not None:
                set_obstacle(self._h, b, kw["circle"])
        self.iteration = 0
        self._tiles = (4, 4, 4, 4)  # refine lanes, depth, gather lanes, depth (the engine's default)
        pos = np.ascontiguousarray(np.concatenate([np.asarray(c[0], np.float32).reshape(-1, 2) for c in cases]))
        f = np.ascontiguousarray(np.concatenate([np.asarray(c[1], np.float32).reshape(-1, 7) for c in cases]))
        counts = np.array([len(np.asarray(c[0]).reshape(-1, 2)) for c in cases], np.int64)
        call("drb_sim_set_state", self._h, pos.ctypes.data, f.ctypes.data, len(pos), counts.ctypes.data)

    def counts(self):
        m = C.c_int64(0)
        out = np.zeros(self.n, np.int64)
        call("drb_sim_count", self._h, C.byref(m), out.ctypes.data)
        return out

    def count(self):
        return int(self.counts().sum())

    def step(self, nsteps: int = 1):
        call("drb_sim_step", self._h, int(nsteps))
        self.iteration += int(nsteps)

    @property
    def updates(self) -> int:
        """Point-updates performed so far, over every instance: the moments entering each step,
        summed over steps."""
        v = C.c_int64(0)
        call("drb_sim_updates", self._h, C.byref(v))
        return int(v.value)

    def moments(self):
        """List of (pos (M_b, 2), f (M_b, 7)) per instance."""
        counts = self.counts()
        m = int(counts.sum())
        pos = np.empty((m, 2), np.float32)
        f = np.empty((m, 7), np.float32)
        if m:
            call("drb_sim_get_state", self._h, pos.ctypes.data, f.ctypes.data)
        offs = np.concatenate([[0], np.cumsum(counts)])
        return [(pos[offs[b]:offs[b + 1]], f[offs[b]:offs[b + 1]]) for b in range(self.n)]

    def save(self, path):
        """Write a checkpoint (.npz) of every instance."""
        from .checkpoint import save

        parts = self.moments()
        pos = np.concatenate([p for p, _ in parts]) if parts else np.zeros((0, 2), np.float32)
        f = np.concatenate([g for _, g in parts]) if parts else np.zeros((0, 7), np.float32)
        t = self._tiles
        save(path, pos, f, self.iteration, self._kws, [len(p) for p, _ in parts], (t[0], t[2]), (t[1], t[3]))

    @classmethod
    def load(cls, path, device=None):
        from .checkpoint import load

        pos, f, iteration, kws, counts, lanes, depth = load(path)
        offs = np.concatenate([[0], np.cumsum(counts)])
        sim = cls([(pos[offs[b]:offs[b + 1]], f[offs[b]:offs[b + 1]], kws[b]) for b in range(len(kws))], device)
        sim.set_tiles(lanes[0], depth[0], lanes[1], depth[1])
        sim.iteration = iteration
        return sim

    def diagnostics(self):
        out = np.zeros(5 * self.n, np.float64)
        call("drb_sim_diagnostics", self._h, out.ctypes.data)
        return out.reshape(self.n, 5)

    def set_timing(self, on=True):
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
        """Step through captured HIP graphs (see GpuSimulator.set_graph)."""
        if horizon is not None:
            call("drb_sim_set_graph_horizon", self._h, int(horizon))
        call("drb_sim_set_graph", self._h, 1 if on else 0)

    def graph_info(self):
        """See GpuSimulator.graph_info."""
        out = (C.c_int64 * 4)()
        call("drb_sim_graph_info", self._h, out)
        return {"captures": int(out[0]), "resumes": int(out[1]), "cells": int(out[2]), "points": int(out[3])}

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            call("drb_sim_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["BatchSimulator"]
