"""Differentiable DiReBM: the GPU step and its hand-written adjoint behind torch.autograd.

No reference counterpart (the reference simulators, direbm/warp/simulator.py and
direbm/reference/simulator.py, are not differentiated).  ``rollout(pos0, f0, tau, steps)`` runs
``steps`` engine steps and returns the final state (pos, f); its backward pass

* restores the checkpoint at the start of each segment of ``checkpoint_every`` steps (the forward
  keeps only those states),
* recomputes the segment with the engine recording every step's intermediates (drb_sim_tape),
* runs the adjoint kernels of the segment's steps in reverse (drb_sim_tape_vjp).

Engines are cached per (device, parameters) so their tape and scratch buffers are reused.

Per step the discrete choices (thinning runs, refine hard/soft, pairs within the radii, empty
directions) are held at the values the forward made; every continuous dependence -- collide,
dispersion shifts, centroids, the exp-weighted refine, the scatter weights and normalisers, the
gap-fill sums -- is differentiated.  Gradients flow to the initial positions, the initial
distributions and tau.  ``equilibrium_t`` makes the initial f differentiable w.r.t. (rho, u).
"""

from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._device import require_gpu, stream_ptr, torch
from ._lib import call
from .simulator import make_params


class _Engine:
    """One drb_sim handle driven with device-resident state."""

    def __init__(self, dev, tau, kw):
        self.dev = dev
        self.params = make_params(tau, **kw)
        self.h = C.c_void_p()
        call("drb_sim_create", C.byref(self.h), C.byref(self.params), 1, stream_ptr(dev))

    def set_state(self, pos, f):
        call("drb_sim_set_state_device", self.h, pos.data_ptr(), f.data_ptr(), pos.shape[0], None)

    def count(self):
        m = C.c_int64(0)
        call("drb_sim_count", self.h, C.byref(m), None)
        return int(m.value)

    def get_state(self):
        t = torch()
        m = self.count()
        pos = t.empty((m, 2), dtype=t.float32, device=self.dev)
        f = t.empty((m, 7), dtype=t.float32, device=self.dev)
        if m:
            call("drb_sim_get_state_device", self.h, pos.data_ptr(), f.data_ptr())
        return pos, f

    def step(self, n):
        call("drb_sim_step", self.h, int(n))

    def tape(self, n):
        call("drb_sim_tape", self.h, int(n))

    def tape_info(self, slot):
        m = C.c_int64(0)
        p = C.c_int64(0)
        call("drb_sim_tape_info", self.h, int(slot), C.byref(m), C.byref(p))
        return int(m.value), int(p.value)

    def step_vjp(self, slot, gpos, gf, gtau):
        """Cotangents of the input of the recorded step `slot` from those of its output."""
        t = torch()
        m, p = self.tape_info(slot)
        if gf.shape[0] != p:
            raise ValueError(f"output cotangent has {gf.shape[0]} rows, the recorded step emitted {p}")
        gpos_in = t.empty((m, 2), dtype=t.float32, device=self.dev)
        gf_in = t.empty((m, 7), dtype=t.float32, device=self.dev)
        call("drb_sim_step_vjp", self.h, int(slot), gpos.data_ptr() if gpos is not None else None, gf.data_ptr(),
             gpos_in.data_ptr(), gf_in.data_ptr(), gtau.ctypes.data)
        return gpos_in, gf_in

    def tape_vjp(self, m_in, gpos, gf, gtau):
        """Cotangents of the input of the first recorded step from those of the last one's output."""
        t = torch()
        gpos_in = t.empty((m_in, 2), dtype=t.float32, device=self.dev)
        gf_in = t.empty((m_in, 7), dtype=t.float32, device=self.dev)
        call("drb_sim_tape_vjp", self.h, gpos.data_ptr() if gpos is not None else None, gf.data_ptr(),
             gpos_in.data_ptr(), gf_in.data_ptr(), gtau.ctypes.data)
        return gpos_in, gf_in

    def close(self):
        if self.h.value:
            call("drb_sim_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FrozenRollout:
    """A rollout recorded step by step, re-evaluated at other initial states with every step's
    recorded discrete choices held fixed (drb_sim_replay): the function whose gradient ``rollout``
    returns.  Central differences of it check that gradient without the jumps a discrete choice
    makes.  Holds one tape slot per step (about 250 bytes per point per step of device memory).

    FrozenRollout(pos0, f0, tau, steps, **kw); ``.final`` = the recorded final (pos, f);
    ``__call__(pos0, f0, tau=None)`` -> (pos, f) after the replayed steps.  float64 inputs replay
    every step in FP64 (drb_sim_replay64: the same function, free of fp32 rounding, so central
    differences of it are accurate at small steps); float32 inputs replay in the engine's fp32.
    """

    def __init__(self, pos0, f0, tau, steps, **kw):
        t = torch()
        dev = require_gpu(pos0.device)
        self.dev, self.steps, self.tau = dev, int(steps), float(tau)
        self.eng = _Engine(dev, float(tau), kw)
        self.eng.set_state(pos0.detach().float().contiguous(), f0.detach().float().contiguous())
        self.eng.tape(self.steps)
        self.eng.step(self.steps)
        self.sizes = [self.eng.tape_info(k) for k in range(self.steps)]
        self.final = self.eng.get_state()
        self._t = t

    def __call__(self, pos0, f0, tau=None):
        t = self._t
        fp64 = f0.dtype == t.float64
        dt = t.float64 if fp64 else t.float32
        fn = "drb_sim_replay64" if fp64 else "drb_sim_replay"
        pos = pos0.detach().to(dt).contiguous()
        f = f0.detach().to(dt).contiguous()
        for k, (m, p) in enumerate(self.sizes):
            po = t.empty((p, 2), dtype=dt, device=self.dev)
            fo = t.empty((p, 7), dtype=dt, device=self.dev)
            call(fn, self.eng.h, k, float(tau) if tau is not None else 0.0, pos.data_ptr(), f.data_ptr(),
                 po.data_ptr(), fo.data_ptr())
            pos, f = po, fo
        return pos, f

    def close(self):
        self.eng.close()


_ENGINES = {}
_MAX_ENGINES = 4


def _engine(dev, tau, kw):
    """An engine for (device, parameters), reused across rollouts: its tape and scratch buffers
    stay allocated, so repeated forward/backward passes allocate nothing.  Every use loads its
    state explicitly, so interleaved rollouts with equal parameters can share it."""
    p = make_params(tau, **kw)
    key = (str(dev),) + tuple(getattr(p, name) for name, _ in p._fields_)
    eng = _ENGINES.pop(key, None)
    if eng is None:
        eng = _Engine(dev, tau, kw)
        while len(_ENGINES) >= _MAX_ENGINES:
            _ENGINES.pop(next(iter(_ENGINES))).close()
    _ENGINES[key] = eng  # most recently used last
    return eng


def _rollout_fn():
    t = torch()

    class Rollout(t.autograd.Function):
        @staticmethod
        def forward(ctx, pos0, f0, tau, steps, every, kw, timings):
            dev = pos0.device
            eng = _engine(dev, float(tau), kw)
            eng.set_state(pos0.detach().contiguous(), f0.detach().contiguous())
            ckpts = [(pos0.detach().clone(), f0.detach().clone())]
            done = 0
            while done < steps:
                n = min(every, steps - done)
                eng.step(n)
                done += n
                if done < steps:
                    ckpts.append(eng.get_state())
            pos, f = eng.get_state()
            ctx.eng, ctx.ckpts, ctx.steps, ctx.every, ctx.timings = eng, ckpts, steps, every, timings
            return pos, f

        @staticmethod
        def backward(ctx, gpos, gf):
            eng, ckpts, steps, every, timings = ctx.eng, ctx.ckpts, ctx.steps, ctx.every, ctx.timings
            dev = eng.dev
            gf = (gf if gf is not None else t.zeros((eng.count(), 7), device=dev)).to(t.float32).contiguous()
            gpos = gpos.to(t.float32).contiguous() if gpos is not None else None
            gtau = np.zeros(1, np.float64)
            nseg = len(ckpts)
            for seg in range(nseg - 1, -1, -1):
                n = min(every, steps - seg * every)
                t0 = time.perf_counter()
                eng.set_state(*ckpts[seg])
                eng.tape(n)
                eng.step(n)
                if timings is not None:
                    t.cuda.synchronize(dev)
                    timings["recompute_s"] = timings.get("recompute_s", 0.0) + time.perf_counter() - t0
                t0 = time.perf_counter()
                gpos, gf = eng.tape_vjp(ckpts[seg][0].shape[0], gpos, gf, gtau)
                if timings is not None:
                    t.cuda.synchronize(dev)
                    timings["backward_s"] = timings.get("backward_s", 0.0) + time.perf_counter() - t0
            ctx.eng = None
            ctx.ckpts = None
            return gpos, gf, t.tensor(gtau[0], dtype=t.float64, device=dev), None, None, None, None

    return Rollout


_ROLLOUT = None


def rollout(pos0, f0, tau, steps, *, checkpoint_every=20, timings=None, **kw):
    """Run `steps` DiReBM steps from (pos0 (M, 2), f0 (M, 7)) and return the final (pos, f).

    Differentiable w.r.t. pos0, f0 and tau (a float or a 0-d tensor) through the engine's adjoint;
    kw are GpuSimulator keywords (dx, alpha, kappa_hard, rho_rest, r_res, box, two_level, eps).
    The forward keeps one state per `checkpoint_every` steps; the backward recomputes each
    segment once.  `timings` (a dict) receives recompute_s / backward_s when given.
    """
    global _ROLLOUT
    t = torch()
    dev = require_gpu(pos0.device if isinstance(pos0, t.Tensor) else None)
    if _ROLLOUT is None:
        _ROLLOUT = _rollout_fn()
    pos0 = t.as_tensor(pos0, dtype=t.float32, device=dev).reshape(-1, 2).contiguous()
    f0 = t.as_tensor(f0, dtype=t.float32, device=dev).reshape(-1, 7).contiguous()
    if not isinstance(tau, t.Tensor):
        tau = t.tensor(float(tau), dtype=t.float64, device=dev)
    if int(checkpoint_every) < 1 or int(steps) < 0:
        raise ValueError("steps >= 0 and checkpoint_every >= 1 required")
    return _ROLLOUT.apply(pos0, f0, tau, int(steps), int(checkpoint_every), dict(kw), timings)


_EQ = None


def _eq_fn():
    t = torch()

    class Equilibrium(t.autograd.Function):
        @staticmethod
        def forward(ctx, rho, u):
            dev = rho.device
            out = t.empty((rho.shape[0], 7), dtype=t.float32, device=dev)
            call("drb_equilibrium", rho.data_ptr(), u.data_ptr(), rho.shape[0], out.data_ptr(), 7, 1, stream_ptr(dev))
            ctx.save_for_backward(rho, u)
            return out

        @staticmethod
        def backward(ctx, g):
            rho, u = ctx.saved_tensors
            dev = rho.device
            g = g.to(t.float32).contiguous()
            grho = t.empty_like(rho)
            gu = t.empty_like(u)
            call("drb_equilibrium_vjp", rho.data_ptr(), u.data_ptr(), rho.shape[0], g.data_ptr(), 7, 1,
                 grho.data_ptr(), gu.data_ptr(), stream_ptr(dev))
            return grho, gu

    return Equilibrium


def equilibrium_t(rho, u):
    """f_eq (N, 7) float32 of rho (N,), u (N, 2) on the GPU, differentiable w.r.t. rho and u."""
    global _EQ
    t = torch()
    if _EQ is None:
        _EQ = _eq_fn()
    return _EQ.apply(rho.to(t.float32).contiguous(), u.to(t.float32).contiguous())


def equilibrium64(rho, u):
    """f_eq (N, 7) in float64 of rho (N,), u (N, 2): the function equilibrium_t evaluates in fp32
    (the engine's fp32 velocity set and weights, cs^2 = 1/4), for central differences in FP64."""
    t = torch()
    s = np.float32(np.sqrt(3.0) / 2.0)
    c = np.array([[0, 0], [1, 0], [0.5, s], [-0.5, s], [-1, 0], [-0.5, -s], [0.5, -s]], np.float32)
    w = np.array([0.5] + [1.0 / 12.0] * 6, np.float32)
    c = t.tensor(c.astype(np.float64), device=u.device)
    w = t.tensor(w.astype(np.float64), device=u.device)
    rho = rho.to(t.float64)
    u = u.to(t.float64)
    cu = u @ c.T
    usq = (u * u).sum(1, keepdim=True)
    return rho[:, None] * w * (1.0 + 4.0 * cu + 8.0 * cu * cu - 2.0 * usq)


def fields_loss(pos, f, edges, h, rho_t, u_t):
    """Mean over bins of (rho - rho_t)^2 + |u - u_t|^2 for the fields binned from (pos, f) with
    bin edges `edges` and cell size h (direbm/fields.py conventions); differentiable w.r.t. f."""
    from .fields import bin_moments, fields_t

    mass, mom = bin_moments(pos, f, edges)
    rho, u = fields_t(mass, mom, h)
    return ((rho - rho_t) ** 2).mean() + ((u - u_t) ** 2).sum(-1).mean()


def fixed_bins(pos, edges):
    """Bin index (iy * nb + ix) of each point for edges `edges` (the rule of bin_moments: a
    coordinate's bin is the number of edges <= it minus one, the last edge in the last bin), -1
    outside; a torch int64 tensor on pos's device."""
    t = torch()
    e = t.as_tensor(np.ascontiguousarray(edges, np.float64), device=pos.device)
    nb = e.shape[0] - 1
    p = pos.detach().double()
    ib = t.searchsorted(e, p.contiguous(), right=True) - 1
    ib = t.where(p == e[-1], t.full_like(ib, nb - 1), ib)
    inside = ((ib >= 0) & (ib < nb)).all(1)
    return t.where(inside, ib[:, 1] * nb + ib[:, 0], t.full_like(ib[:, 0], -1))


def fields_loss_fixed(bins, f, nb, h, rho_t, u_t):
    """fields_loss with each point's bin given (fixed_bins) and every operation in f's precision
    (FP64 for central differences of drb_sim_replay64): the same function as fields_loss at the
    positions the bins were taken from."""
    t = torch()
    keep = bins >= 0
    b = bins[keep]
    ff = f[keep]
    c = t.tensor([[0.0, 0.0], [1.0, 0.0], [0.5, 0.8660254037844386], [-0.5, 0.8660254037844386], [-1.0, 0.0],
                  [-0.5, -0.8660254037844386], [0.5, -0.8660254037844386]], dtype=f.dtype, device=f.device)
    mass = t.zeros(nb * nb, dtype=f.dtype, device=f.device).index_add_(0, b, ff.sum(1))
    mom = t.zeros((nb * nb, 2), dtype=f.dtype, device=f.device).index_add_(0, b, ff @ c)
    from .fields import fields_t

    rho, u = fields_t(mass.reshape(nb, nb), mom.reshape(nb, nb, 2), h)
    return ((rho - rho_t.to(f.dtype)) ** 2).mean() + ((u - u_t.to(f.dtype)) ** 2).sum(-1).mean()


__all__ = ["rollout", "equilibrium_t", "equilibrium64", "fields_loss", "fields_loss_fixed", "fixed_bins",
           "FrozenRollout"]
