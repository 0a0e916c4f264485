"""The FP64 autograd reference of the step adjoint (oracle/adjoint.py), pinned against v2.step:

* its forward value equals v2.step's output (choices fixed at the state's own values);
* its autograd derivative equals central differences of v2.step itself (choices recomputed at
  every evaluation) along random directions in f, tau and positions, where no choice flips;
* collide_t's adjoint satisfies the collision invariants (mass and momentum cotangents pass
  through unchanged, no tau dependence).
"""

import numpy as np
import pytest
import torch

from oracle import adjoint, v2
from oracle.physics import collide, equilibrium
from tests._util import random_cloud


def _pulse_state(n=6, seed=0):
    """Jittered grid with a density bump, open domain, after 2 oracle steps (generic state)."""
    rng = np.random.default_rng(seed)
    g = np.arange(n, dtype=np.float64) * 0.25
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.02, 0.02, (n * n, 2))
    r2 = ((pos - pos.mean(0)) ** 2).sum(1)
    rho = (1 + 0.05 * np.exp(-r2 / 0.2)) * 0.0625
    u = rng.uniform(-0.02, 0.02, (n * n, 2))
    f = equilibrium(rho, u)
    prm = v2.Params(tau=0.7, dx=1.0, alpha=4.0, rho_rest=0.0625)
    for _ in range(2):
        r = v2.step(pos, f, prm)
        pos, f = r.pos.astype(np.float32), r.f.astype(np.float32)
    return pos.astype(np.float32), f.astype(np.float32), prm


def _tg_state(n=16, seed=1):
    L = 2 * np.pi
    h = L / n
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))
    u = np.stack([-0.05 * np.cos(p[:, 0]) * np.sin(p[:, 1]), 0.05 * np.sin(p[:, 0]) * np.cos(p[:, 1])], 1)
    f = equilibrium(np.full(n * n, h * h), u)
    prm = v2.Params(tau=0.8, dx=4 * h, alpha=4.0, rho_rest=h * h, box=L)
    r = v2.step(p.astype(np.float32), f, prm)
    return r.pos.astype(np.float32), r.f.astype(np.float32), prm


CASES = {"open_pulse": _pulse_state, "periodic_tg": _tg_state}


@pytest.mark.parametrize("name", sorted(CASES))
def test_forward_equals_v2_step(name):
    pos, f, prm = CASES[name]()
    fn = adjoint.StepFn(pos, f, prm)
    with torch.no_grad():
        x, out = fn(torch.tensor(pos, dtype=torch.float64), torch.tensor(f, dtype=torch.float64),
                    torch.tensor(prm.tau, dtype=torch.float64))
    ref = v2.step(pos, f, prm)
    assert x.shape == ref.pos.shape and out.shape == ref.f.shape
    assert np.array_equal(x.numpy(), ref.pos)
    assert np.allclose(out.numpy(), ref.f, rtol=1e-12, atol=1e-18)


def _loss_v2(pos, f, prm, gpos, gf):
    r = v2.step(pos, f, prm, round_positions=False)
    if len(r.pos) != len(gpos):  # a control point appeared or vanished: a discrete change
        return None, r
    return float((r.pos * gpos).sum() + (r.f * gf).sum()), r


def _same_choices(ra, rb):
    return (len(ra.pos) == len(rb.pos) and np.array_equal(ra.thin.order, rb.thin.order)
            and np.array_equal(ra.thin.start, rb.thin.start) and np.array_equal(ra.kappa, rb.kappa)
            and np.array_equal(ra.res.cp_f < 1e-6, rb.res.cp_f < 1e-6))


def _pairs(pos, f, prm):
    """Every pair set of the step (refine disks, scatter pairs, gap-fill pairs): the refine
    weights exp(v) do not vanish at the disk edge, so a pair entering it is a discrete change."""
    fn = adjoint.StepFn(pos, f, prm, round_positions=False)
    return [np.stack(sorted(zip(a.tolist(), b.tolist()))) if len(a) else np.zeros((0, 2))
            for a, b in ((fn.ref_ri, fn.ref_ci), (fn.sc, fn.sp), (fn.gi, fn.gj))]


def _same_pairs(pa, pb):
    return all(a.shape == b.shape and np.array_equal(a, b) for a, b in zip(pa, pb))


@pytest.mark.parametrize("name", sorted(CASES))
def test_autograd_matches_central_differences_of_v2(name):
    pos, f, prm = CASES[name]()
    rng = np.random.default_rng(7)
    P = len(v2.step(pos, f, prm).pos)
    gpos = rng.standard_normal((P, 2))
    gf = rng.standard_normal((P, 7))
    dpos, df, dtau, _ = adjoint.step_vjp(pos, f, prm, gpos, gf)
    # the fp32 rounding of the refined positions only moves the evaluation point (by an ulp)
    dpos2, df2, dtau2, _ = adjoint.step_vjp(pos, f, prm, gpos, gf, round_positions=False)
    assert np.linalg.norm(dpos - dpos2) <= 1e-5 * np.linalg.norm(dpos)
    assert np.linalg.norm(df - df2) <= 1e-5 * np.linalg.norm(df)
    assert abs(dtau - dtau2) <= 1e-5 * abs(dtau)
    base_loss, base = _loss_v2(pos, f, prm, gpos, gf)

    # f and tau: differences of v2.step with the refined positions kept in FP64 (the fp32 store
    # would turn small position changes into steps); relative perturbations keep small values
    # away from the emptiness threshold; directions that flip a choice are skipped
    ok = 0
    for trial in range(6):
        d = rng.standard_normal(f.shape) * np.abs(f)
        eps = 1e-3
        lp, rp = _loss_v2(pos, f.astype(np.float64) + eps * d, prm, gpos, gf)
        lm, rm = _loss_v2(pos, f.astype(np.float64) - eps * d, prm, gpos, gf)
        if not (_same_choices(base, rp) and _same_choices(base, rm)):
            continue
        fd = (lp - lm) / (2 * eps)
        an = float((df * d).sum())
        assert abs(fd - an) <= 2e-3 * abs(an), (trial, fd, an)
        ok += 1
    assert ok >= 2

    # tau
    eps = 1e-4
    lp, rp = _loss_v2(pos, f, v2.Params(**{**prm.__dict__, "tau": prm.tau + eps}), gpos, gf)
    lm, rm = _loss_v2(pos, f, v2.Params(**{**prm.__dict__, "tau": prm.tau - eps}), gpos, gf)
    assert _same_choices(base, rp) and _same_choices(base, rm)
    fd = (lp - lm) / (2 * eps)
    assert abs(fd - dtau) <= 2e-3 * abs(dtau), (fd, dtau)

    # positions: v2.step takes them in fp32, so the step must be resolvable in fp32; directions
    # whose perturbation flips a discrete choice are skipped (at least two must remain)
    ok = 0
    base_pairs = _pairs(pos, f, prm)
    for trial in range(24):
        d = np.zeros(pos.shape)
        d[rng.integers(len(pos))] = rng.standard_normal(2)  # one point at a time: fewer flips
        eps = 2e-4 * prm.dx
        pp = (pos.astype(np.float64) + eps * d).astype(np.float32)
        pm = (pos.astype(np.float64) - eps * d).astype(np.float32)
        lp, rp = _loss_v2(pp, f, prm, gpos, gf)
        lm, rm = _loss_v2(pm, f, prm, gpos, gf)
        if lp is None or lm is None or not (_same_choices(base, rp) and _same_choices(base, rm)):
            continue
        if not (_same_pairs(base_pairs, _pairs(pp, f, prm)) and _same_pairs(base_pairs, _pairs(pm, f, prm))):
            continue
        delta = pp.astype(np.float64) - pm.astype(np.float64)  # the perturbation fp32 really made
        fd = lp - lm
        an = float((dpos * delta).sum())
        # fp32 rounding of the component positions and centroids: ~1e-7 against steps of 2e-4
        assert abs(fd - an) <= 1e-2 * abs(an) + 1e-9, (trial, fd, an)
        ok += 1
        if ok == 3:
            break
    assert ok >= 3


def test_collide_adjoint_invariants():
    rng = np.random.default_rng(3)
    f = torch.tensor(equilibrium(rng.uniform(0.5, 2, 50), rng.uniform(-0.1, 0.1, (50, 2)))
                     + rng.uniform(0, 0.01, (50, 7)), requires_grad=True)
    tau = torch.tensor(0.7, dtype=torch.float64, requires_grad=True)
    out = adjoint.collide_t(f, tau)
    assert np.allclose(out.detach().numpy(), collide(f.detach().numpy(), 0.7), rtol=1e-14)
    from oracle.constants import C2
    for g in (np.ones(7), C2[:, 0], C2[:, 1]):
        gf, gt = torch.autograd.grad(out, (f, tau), torch.tensor(np.tile(g, (50, 1))), retain_graph=True)
        assert np.allclose(gf.numpy(), np.tile(g, (50, 1)), atol=1e-12)
        assert abs(float(gt)) < 1e-12


def test_random_cloud_vjp_runs():
    pos, f = random_cloud(40, seed=5)
    prm = v2.Params(tau=0.6, dx=1.0, alpha=4.0, rho_rest=0.1)
    r = v2.step(pos, f, prm)
    dpos, df, dtau, _ = adjoint.step_vjp(pos, f, prm, np.ones_like(r.pos), np.ones_like(r.f))
    assert np.isfinite(dpos).all() and np.isfinite(df).all() and np.isfinite(dtau)
