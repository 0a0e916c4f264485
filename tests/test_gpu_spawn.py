"""The soft_outer spawn on the GPU (kappa_soft; direbm/reference/simulator.py:141-166 carried into the
cell-thinning step): teacher-forced steps against the FP64 oracle (same point count, spawned
positions bit-equal, f relative L2 <= 1e-5), bit-identity with the plain step when it is off, the
reference's behaviour test on the engine (the spawn makes a point pulse's front more isotropic,
tests/test_reference.py:70-85), the mass budget, and the modes that refuse it."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from direbm_amd._lib import DrbError
from oracle import v2
from oracle.physics import equilibrium
from tests._util import min_image, random_cloud, rel_l2

pytestmark = pytest.mark.gpu

KS = 5  # KAPPA_SOFT of direbm/constants.py


def _point_pulse(rho=1.6):
    return np.zeros((1, 2), np.float32), equilibrium(np.float64(rho), np.zeros(2))[None].astype(np.float32)


def _cases():
    pos, f = workloads.rest_pulse(5, pulse_rho=1.5)
    yield "rest_pulse", pos.astype(np.float32), f.astype(np.float32), dict(tau=0.6)
    p, f = _point_pulse()
    yield "point_pulse", p, f, dict(tau=0.6)
    g = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    yield "two_level", g.pos, g.f, dict(g.kw)
    c = workloads.taylor_green(n=24, tau=0.7)
    yield "periodic_tg", c.pos, c.f, dict(c.kw)
    pm, fm = random_cloud(300, 7, -6.0, 6.0)
    yield "random_cloud", pm.astype(np.float32), (0.2 * fm).astype(np.float32), dict(tau=0.7)


def _oracle_params(kw):
    return v2.Params(tau=kw["tau"], dx=kw.get("dx", 1.0), alpha=kw.get("alpha", 4.0), rho_rest=kw.get("rho_rest", 1.0),
                     box=kw.get("box"), two_level=kw.get("two_level"), kappa_soft=KS)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_spawn_teacher_forced_against_oracle(case):
    name, pos, f, kw = case
    sim = drb.GpuSimulator(pos, f, kappa_soft=KS, **kw)
    prm = _oracle_params(kw)
    spawned = 0
    for step in range(3):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        n_reg = len(ref.thin.cp32)
        spawned += len(ref.pos) - n_reg
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos), f"{name} step {step}: point count"
        # the spawned points (after the refined ones) are formed in FP64 and rounded once: bit-equal
        assert np.array_equal(p1[n_reg:], ref.pos[n_reg:].astype(np.float32)), f"{name} step {step}: spawned"
        d = np.abs(min_image(p1.astype(np.float64) - ref.pos, kw.get("box")))
        assert d.max() <= 1e-5 * max(1.0, np.abs(ref.pos).max()), f"{name} step {step}: positions"
        e = rel_l2(f1, ref.f)
        assert e <= 1e-5, f"{name} step {step}: relative L2 {e:.3g}"
    if name != "periodic_tg":  # a saturated periodic field has no soft points
        assert spawned > 0, name


def test_spawn_off_is_the_plain_step():
    g = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(g.pos, g.f, **g.kw)
    b = drb.GpuSimulator(g.pos, g.f, kappa_soft=4, **g.kw)  # kappa_soft <= kappa_hard: off
    a.step(3)
    b.step(3)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def _front_anisotropy(kappa_soft, steps=10, sectors=36):
    p, f = _point_pulse()
    sim = drb.GpuSimulator(p, f, tau=0.6, kappa_soft=kappa_soft)
    sim.step(steps)
    x, _ = sim.moments()
    r = np.linalg.norm(x.astype(np.float64), axis=1)
    th = np.arctan2(x[:, 1], x[:, 0]) % (2 * np.pi)
    edges = np.linspace(0, 2 * np.pi, sectors + 1)
    front = np.array([r[(th >= edges[k]) & (th < edges[k + 1])].max(initial=0.0) for k in range(sectors)])
    return front.std() / front.mean()


def test_spawn_reduces_circular_anisotropy_like_reference():
    # tests/test_reference.py:82-85 on the engine
    on, off = _front_anisotropy(KS), _front_anisotropy(None)
    print(f"front anisotropy after 10 steps: spawn {on:.4f}, off {off:.4f}", flush=True)
    assert on < 0.5 * off


def test_spawn_mass_budget_and_determinism():
    g = workloads.gaussian_pulse_two_level(n=96, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(g.pos, g.f, kappa_soft=KS, **g.kw)
    b = drb.GpuSimulator(g.pos, g.f, kappa_soft=KS, **g.kw)
    m_prev = a.diagnostics()["mass"]
    for _ in range(5):
        a.step()
        b.step()
        d = a.diagnostics()
        assert abs(d["mass"] - (m_prev - d["dropped"] + d["injected"])) <= 1e-6 * m_prev
        m_prev = d["mass"]
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def test_spawn_refusals():
    p, f = _point_pulse()
    with pytest.raises(DrbError, match="INVALID_ARG"):  # batches
        drb.BatchSimulator([(p, f, dict(tau=0.6, kappa_soft=KS)), (p, f, dict(tau=0.6))])
    with pytest.raises(DrbError, match="INVALID_ARG"):  # with an obstacle
        drb.GpuSimulator(p + 5.0, f, tau=0.6, kappa_soft=KS, circle=(0.0, 0.0, 1.0))
    c = workloads.taylor_green(n=16)
    sim = drb.GpuSimulator(c.pos, c.f, kappa_soft=KS, **c.kw)
    with pytest.raises(DrbError, match="INVALID_ARG"):  # graph mode
        sim.set_graph(True)
    c3 = workloads.pulse_3d(n=4)
    with pytest.raises(DrbError, match="INVALID_ARG"):  # 3D (the reference's spawn is 2D only)
        drb.GpuSimulator3D(c3.pos, c3.f, kappa_soft=KS, **c3.kw)
