"""FP64 autograd reference of one DiReBM step with its discrete choices held fixed (test
infrastructure: the checker of the engine's hand-written adjoint, never the product path).

The step is oracle/v2.py's restatement of direbm/warp/simulator.py:50-73.  ``StepFn(pos, f, prm)``
evaluates every discrete decision of that step from the fp32 state exactly as v2.step does (cell
keys and runs, refine pairs and kappa, scatter pairs, gap-fill pairs and empty directions) and
then exposes the step as a smooth torch float64 function of (pos, f, tau) with those choices
fixed.  Autograd of it is the reference VJP; its forward value equals v2.step's (pinned by
tests/test_oracle_adjoint.py, which also checks it against central differences of v2.step
itself, choices recomputed, where no choice flips).
"""

from __future__ import annotations

import numpy as np
import torch

from . import v2
from .constants import C2, CS2_2D, W2

F32 = np.float32


def collide_t(f, tau):
    """BGK collide (direbm/physics.py:48-55) in torch float64, rho == 0 -> u = 0."""
    C = torch.as_tensor(C2, dtype=torch.float64)
    W = torch.as_tensor(W2, dtype=torch.float64)
    rho = f.sum(1)
    j = f @ C
    live = rho != 0
    u = torch.where(live[:, None], j / torch.where(live, rho, torch.ones_like(rho))[:, None], torch.zeros_like(j))
    inv = 1.0 / CS2_2D
    cu = u @ C.T
    usq = (u * u).sum(1, keepdim=True)
    feq = rho[:, None] * W[None, :] * (1 + inv * cu + 0.5 * inv * inv * cu * cu - 0.5 * inv * usq)
    return f + (feq - f) / tau


def equilibrium_t(rho, u):
    """f_eq (direbm/physics.py:29-45) in torch float64."""
    C = torch.as_tensor(C2, dtype=torch.float64)
    W = torch.as_tensor(W2, dtype=torch.float64)
    inv = 1.0 / CS2_2D
    cu = u @ C.T
    usq = (u * u).sum(1, keepdim=True)
    return rho[:, None] * W[None, :] * (1 + inv * cu + 0.5 * inv * inv * cu * cu - 0.5 * inv * usq)


def _norm(d):
    """Row norms with gradient 0 where the norm is 0 (the cone R - |d| has no derivative there;
    the engine's adjoint also skips zero-distance pairs)."""
    zero = (d == 0).all(1)
    n = torch.sqrt(torch.where(zero[:, None], torch.ones_like(d), d * d).sum(1))
    return torch.where(zero, torch.zeros_like(n), n)


def _mi_shift(d64, L):
    """Shift of the minimum-image convention for FP64 differences of fp32 coordinates."""
    if L is None:
        return np.zeros_like(d64)
    L64 = float(F32(L))
    return np.where(d64 > 0.5 * L64, -L64, np.where(d64 < -0.5 * L64, L64, 0.0))


class StepFn:
    """One step from the fp32 state (pos (M, 2), f (M, 7)) with its choices fixed.

    ``StepFn(pos32, f32, prm)(pos, f, tau) -> (pos', f')`` for torch float64 (pos, f, tau).
    """

    def __init__(self, pos32, f32, prm: v2.Params, round_positions: bool = True):
        self.prm = prm
        pos32 = np.asarray(pos32, F32).reshape(-1, 2)
        f64 = np.asarray(f32, np.float64).reshape(-1, 7)
        self.pos32 = pos32
        self.f64 = f64
        L = prm.box
        self.L = L
        M = len(pos32)
        r = v2.step(pos32, f64, prm, round_positions)
        self.ref = r
        fstar = r.f_star
        pos_c, val_c, dir_c = v2.disperse(pos32, fstar, prm.dx, prm)
        self.dir = dir_c.astype(np.int64)
        parent = np.repeat(np.arange(M), 7)
        self.parent = parent
        # q_c = p_i + C_k dx + s_c: the constant s_c carries the fp32 rounding and the wrap
        self.qconst = pos_c.astype(np.float64) - pos32.astype(np.float64)[parent]
        th = r.thin
        P = len(th.start)
        self.P = P
        run = np.empty(len(pos_c), np.int64)
        for k, (s0, n) in enumerate(zip(th.start, th.length)):
            run[th.order[s0:s0 + n]] = k
        self.run = run
        self.nrun = th.length.astype(np.float64)
        cp32 = th.cp32
        # refine pairs (fp32 predicate at the fp32 centroid)
        ri, ci, _ = v2.pairs_within(cp32, pos_c, prm.dx, L)
        self.ref_ri, self.ref_ci = ri, ci
        self.hard = r.kappa <= prm.kappa_hard
        self.ref_shift = _mi_shift(np.asarray(pos_c[ci], np.float64) - np.asarray(cp32[ri], np.float64), L)
        cp_next = r.pos.astype(F32)
        self.cp_next = r.pos  # value of the refined positions (fp32-rounded unless round_positions=False)
        # scatter pairs
        sc, sp, _ = v2.pairs_within(pos_c, cp_next, prm.rr, L)
        self.sc, self.sp = sc, sp
        self.sc_shift = _mi_shift(np.asarray(cp_next[sp], np.float64) - np.asarray(pos_c[sc], np.float64), L)
        self.empty = r.res.cp_f < prm.eps
        gi, gj, _ = v2.pairs_within(cp_next, cp_next, prm.rr, L)
        self.gi, self.gj = gi, gj
        self.g_shift = _mi_shift(np.asarray(cp_next[gj], np.float64) - np.asarray(cp_next[gi], np.float64), L)
        self.frest = prm.f_rest()
        # the engine stores positions in fp32 (and wraps them): the difference between the stored
        # refined positions and their FP64 value at this state is a constant of the function
        self.xconst = None
        with torch.no_grad():
            x64, _ = self(torch.as_tensor(pos32.astype(np.float64)), torch.as_tensor(f64),
                          torch.tensor(float(prm.tau), dtype=torch.float64))
        self.xconst = torch.as_tensor(np.asarray(r.pos, np.float64)) - x64

    def __call__(self, pos, f, tau):
        prm, L = self.prm, self.L
        T = lambda a: torch.as_tensor(a)  # noqa: E731
        fstar = collide_t(f, tau)
        val = fstar.reshape(-1)
        q = pos[T(self.parent)] + T(self.qconst)
        P = self.P
        # centroid
        a = torch.zeros((P, 2), dtype=torch.float64).index_add(0, T(self.run), q) / T(self.nrun)[:, None]
        # refine
        ri, ci = T(self.ref_ri), T(self.ref_ci)
        w = torch.exp(val[ci])
        d = q[ci] - a[ri] + T(self.ref_shift)
        W = torch.zeros(P, dtype=torch.float64).index_add(0, ri, w)
        S = torch.zeros((P, 2), dtype=torch.float64).index_add(0, ri, w[:, None] * d)
        Wsafe = torch.where(W > 0, W, torch.ones_like(W))
        hard = T(self.hard) | (W <= 0)
        x = torch.where(hard[:, None], a, a + S / Wsafe[:, None])
        if self.xconst is None:  # construction: the FP64 refined positions themselves
            return x, None
        x = x + self.xconst
        # scatter
        sc, sp = T(self.sc), T(self.sp)
        dd = x[sp] - q[sc] + T(self.sc_shift)
        dist = _norm(dd)
        wt = prm.rr - dist
        nc = q.shape[0]
        Sc = torch.zeros(nc, dtype=torch.float64).index_add(0, sc, wt)
        alive = Sc[sc] > 0
        contrib = torch.where(alive, wt / torch.where(alive, Sc[sc], torch.ones_like(wt)) * val[sc],
                              torch.zeros_like(wt))
        F = torch.zeros(P * 7, dtype=torch.float64).index_add(0, sp * 7 + T(self.dir)[sc], contrib).reshape(P, 7)
        # gap-fill
        gi, gj = T(self.gi), T(self.gj)
        gd = x[gj] - x[gi] + T(self.g_shift)
        gw = prm.rr - _norm(gd)  # the self pair: distance 0, weight R
        E = T(self.empty)
        out = F
        for k in range(7):
            G = torch.zeros(P, dtype=torch.float64).index_add(0, gi, gw * E[gj, k].to(torch.float64))
            fill = E[:, k] & (G > 0)
            col = torch.where(fill, prm.rr / torch.where(fill, G, torch.ones_like(G)) * float(self.frest[k]), F[:, k])
            out = torch.cat([out[:, :k], col[:, None], out[:, k + 1:]], 1)
        return x, out


def step_vjp(pos32, f32, prm: v2.Params, gpos, gf, round_positions=True):
    """Reference VJP of one step: (d pos, d f, d tau) for output cotangents (gpos (P,2), gf (P,7))."""
    fn = StepFn(pos32, f32, prm, round_positions)
    pos = torch.tensor(np.asarray(pos32, np.float64), requires_grad=True)
    f = torch.tensor(np.asarray(f32, np.float64), requires_grad=True)
    tau = torch.tensor(float(prm.tau), dtype=torch.float64, requires_grad=True)
    x, out = fn(pos, f, tau)
    loss = (x * torch.as_tensor(np.asarray(gpos, np.float64))).sum() + (out * torch.as_tensor(np.asarray(gf, np.float64))).sum()
    loss.backward()
    return pos.grad.numpy(), f.grad.numpy(), float(tau.grad), fn


class RolloutRef:
    """FP64 composition of StepFn over a recorded trajectory: the rollout with every step's
    discrete choices held at the trajectory's values.  states[t] = (pos32, f32) entering step t."""

    def __init__(self, states, prm: v2.Params):
        self.fns = [StepFn(p, f, prm) for p, f in states]

    def __call__(self, pos, f, tau):
        for fn in self.fns:
            pos, f = fn(pos, f, tau)
        return pos, f


def binned_loss(pos_base, edges, h, rho_t, u_t):
    """Mean over bins of (rho - rho_t)^2 + |u - u_t|^2 (direbm/fields.py conventions) with the bin
    of every moment fixed at pos_base; returns loss(f) for torch float64 f."""
    pos_base = np.asarray(pos_base, np.float64)
    nb = len(edges) - 1
    ix = np.searchsorted(edges, pos_base, side="right") - 1
    ix[pos_base == edges[-1]] -= 1
    inside = ((ix >= 0) & (ix < nb)).all(1)
    b = torch.as_tensor(np.where(inside, ix[:, 1] * nb + ix[:, 0], nb * nb))
    C = torch.as_tensor(C2, dtype=torch.float64)
    rho_t = torch.as_tensor(np.asarray(rho_t, np.float64)).reshape(-1)
    u_t = torch.as_tensor(np.asarray(u_t, np.float64)).reshape(-1, 2)

    def loss(f):
        mass = torch.zeros(nb * nb + 1, dtype=torch.float64).index_add(0, b, f.sum(1))[:-1]
        mom = torch.zeros((nb * nb + 1, 2), dtype=torch.float64).index_add(0, b, f @ C)[:-1]
        empty = mass == 0
        u = mom / torch.where(empty, torch.ones_like(mass), mass)[:, None]
        u = torch.where(empty[:, None], torch.zeros_like(u), u)
        return ((mass / (h * h) - rho_t) ** 2).mean() + ((u - u_t) ** 2).sum(1).mean()

    return loss
