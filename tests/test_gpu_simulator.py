"""Full HIP DiReBM step: reference behaviour tests plus teacher-forced FP64-oracle parity.

Restates tests/test_warp_simulator.py:24-46 and adds, per step and from the same fp32 state:
the same control points (count and order), positions to fp32 rounding, and distribution /
density / momentum relative L2 <= 1e-5 against oracle/v2.py; bit-for-bit determinism; FP64
mass conservation on a saturated periodic field.
"""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle import v2
from oracle.fields import bin_fields
from oracle.physics import equilibrium
from tests._util import min_image, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _oracle_params(kw):
    return v2.Params(tau=kw["tau"], dx=kw.get("dx", 1.0), alpha=kw.get("alpha", 4.0),
                     kappa_hard=kw.get("kappa_hard", 4), rho_rest=kw.get("rho_rest", 1.0),
                     r_res=kw.get("r_res"), box=kw.get("box"), two_level=kw.get("two_level"))


def _teacher_forced(pos, f, kw, nsteps, pointwise=None, refs=None):
    sim = drb.GpuSimulator(pos, f, **kw)
    prm = _oracle_params(kw)
    worst = 0.0
    for _ in range(nsteps):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        if refs is not None:
            refs.append(ref)
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos), "control-point count differs"
        if pointwise is not None:
            e = np.abs(f1.astype(np.float64) - ref.f).max() / np.abs(ref.f).max()
            assert e <= pointwise, f"largest per-point deviation {e:.3g} > {pointwise}"
        d = np.abs(min_image(p1.astype(np.float64) - ref.pos, kw.get("box")))
        assert d.max() <= 1e-5 * max(1.0, np.abs(ref.pos).max())
        rho_g, rho_r = f1.astype(np.float64).sum(1), ref.f.sum(1)
        mom_g, mom_r = f1.astype(np.float64) @ v2.C2, ref.f @ v2.C2
        e = max(rel_l2(f1, ref.f), rel_l2(rho_g, rho_r), rel_l2(mom_g, mom_r) if np.abs(mom_r).max() > 0 else 0.0)
        worst = max(worst, e)
        assert e <= TOL, f"relative L2 {e:.3g} > {TOL}"
    return sim, worst


def test_gpu_rest_field_preserved():
    pos, f = workloads.rest_pulse(6)
    sim = drb.GpuSimulator(pos, f, tau=0.6, rho_rest=1.0)
    for _ in range(4):
        sim.step()
    xp, fp = sim.moments()
    assert np.isfinite(fp).all()
    assert (fp.sum(axis=1) > 0).all()
    rho, _, _ = bin_fields(xp, fp, -3, 3, h=1.0)
    assert abs(rho[rho > 0].mean() - 1.0) < 0.2


def test_gpu_pulse_spreads():
    pos = np.zeros((1, 2))
    f = equilibrium(np.float64(1.5), np.zeros(2))[None, :]
    sim = drb.GpuSimulator(pos, f, tau=0.6)
    for _ in range(5):
        sim.step()
    xp, fp = sim.moments()
    assert len(xp) > 1
    assert np.isfinite(fp).all()
    assert np.linalg.norm(xp, axis=1).max() > 3.0


def test_teacher_forced_rest_pulse():
    pos, f = workloads.rest_pulse(6, pulse_rho=1.5)
    _teacher_forced(pos, f, dict(tau=0.6, rho_rest=1.0), 4)


def test_teacher_forced_single_pulse():
    pos = np.zeros((1, 2))
    f = equilibrium(np.float64(1.6), np.zeros(2))[None, :]
    _teacher_forced(pos, f, dict(tau=0.6, rho_rest=1.0), 6)


def test_teacher_forced_periodic_taylor_green():
    case = workloads.taylor_green(n=32)
    sim, _ = _teacher_forced(case.pos, case.f, case.kw, 3)
    assert sim.count() == 32 * 32


def test_teacher_forced_periodic_seam_crossings_per_point():
    # A blob straddling the periodic seam x = 0 of an 8 x 8 box (dx 1, 32 cells a side): refine
    # moves control points out of their cells and some across the seam, while they stay listed
    # under their pre-refine cells; the normaliser must still find them by the minimum image.
    # Per-point bound, so that one missed point (a local error of 1e-3 or more) cannot hide in a
    # norm over the field.
    rng = np.random.default_rng(2)
    L = 8.0
    g = np.arange(-8, 8) * 0.25 + 0.125
    X, Y = np.meshgrid(g, 3.0 + np.arange(16) * 0.25 + 0.125, indexing="ij")
    pos = (np.stack([X.ravel() % L, Y.ravel()], 1) + rng.uniform(-0.03, 0.03, (256, 2))) % L
    u = np.tile([0.05, 0.02], (256, 1))
    f = equilibrium((1.0 + 0.3 * rng.uniform(size=256)) * 0.0625, u)
    kw = dict(tau=0.7, dx=1.0, alpha=4.0, rho_rest=0.0625, box=L)
    refs = []
    _teacher_forced(pos.astype(np.float32), f.astype(np.float32), kw, 3, pointwise=1e-4, refs=refs)
    crossed = 0
    for ref in refs:
        c0 = np.floor(ref.thin.cp32.astype(np.float64) / 0.25)
        c1 = np.floor(ref.pos / 0.25)
        crossed += int((np.abs(c1 - c0) > 16).any(axis=1).sum())
    assert crossed > 0, "no control point crossed a seam: the case does not exercise the wrap"
    print(f"control points that crossed a seam: {crossed}")


def test_teacher_forced_periodic_resample_radius():
    case = workloads.taylor_green(n=32, r_res=1.5 * 2 * np.pi / 32)
    _teacher_forced(case.pos, case.f, case.kw, 2)


def test_teacher_forced_two_level():
    case = workloads.gaussian_pulse_two_level(n=64, amp=0.05, sigma_cells=4.0)
    _teacher_forced(case.pos, case.f, case.kw, 3)


def test_deterministic_bits():
    case = workloads.taylor_green(n=48)
    a = drb.GpuSimulator(case.pos, case.f, **case.kw)
    b = drb.GpuSimulator(case.pos, case.f, **case.kw)
    a.step(5)
    b.step(5)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def test_periodic_mass_conserved_fp64():
    case = workloads.taylor_green(n=64)
    sim = drb.GpuSimulator(case.pos, case.f, **case.kw)
    m0 = sim.diagnostics()["mass"]
    for _ in range(10):
        sim.step()
        d = sim.diagnostics()
        assert d["dropped"] == 0.0
        assert abs(d["mass"] - m0) <= 1e-6 * m0 + d["injected"]
    assert abs(sim.diagnostics()["mass"] / m0 - 1.0) <= 1e-6


def test_mass_budget_open_domain():
    pos, f = workloads.rest_pulse(5, pulse_rho=1.5)
    sim = drb.GpuSimulator(pos, f, tau=0.6, rho_rest=1.0)
    m_prev = sim.diagnostics()["mass"]
    for _ in range(3):
        sim.step()
        d = sim.diagnostics()
        # m_{t+1} = m_t - dropped + injected (collide conserves mass)
        assert abs(d["mass"] - (m_prev - d["dropped"] + d["injected"])) <= 1e-5 * m_prev
        m_prev = d["mass"]



def test_free_running_agrees_with_oracle_short_horizon():
    """Config-1 shape (periodic Taylor-Green, r_res 1.5h), GPU and FP64 oracle each run free for 10
    steps: same control-point counts every step, Taylor-Green amplitude within 1e-4, mass within
    1e-6, binned fields (bins of 16h) within 2e-3 (density) and 1e-2 (velocity).  Longer horizons are not comparable point-wise: the reference
    algorithm itself amplifies 1-ulp position differences (its own FP64 run perturbed by one ulp
    in 10% of the positions drifts by 1.7e-3 in amplitude after 50 steps)."""
    n = 64
    L = 2 * np.pi
    h = L / n
    c = workloads.taylor_green(n=n, tau=0.8, r_res=1.5 * h)
    kw = c.kw
    sim = drb.GpuSimulator(c.pos, c.f, **kw)
    prm = _oracle_params(kw)
    ps, fs = c.pos.copy(), c.f.copy()

    def amp(p, f):
        p, f = p.astype(np.float64), f.astype(np.float64)
        j, m = f @ v2.C2, f.sum(1)
        bx, by = -np.cos(p[:, 0]) * np.sin(p[:, 1]), np.sin(p[:, 0]) * np.cos(p[:, 1])
        return (j[:, 0] * bx + j[:, 1] * by).sum() / (m * (bx * bx + by * by)).sum()

    for _ in range(10):
        sim.step()
        r = v2.step(ps, fs, prm)
        ps, fs = r.pos.astype(np.float32), r.f.astype(np.float32)
        assert sim.count() == len(ps)
    pg, fg = sim.moments()
    assert abs(amp(pg, fg) - amp(ps, fs)) <= 1e-4 * abs(amp(ps, fs))
    mg, mo = fg.astype(np.float64).sum(), fs.astype(np.float64).sum()
    assert abs(mg - mo) <= 1e-6 * mo
    # binned fields also carry bin-membership noise (one point of 256 crossing a bin edge moves
    # that bin's density by 0.4%): a looser bar
    rg, ug, _ = bin_fields(pg, fg, 0, L, 16 * h)
    ro, uo, _ = bin_fields(ps, fs, 0, L, 16 * h)
    assert rel_l2(rg, ro) <= 2e-3 and rel_l2(ug, uo) <= 1e-2
