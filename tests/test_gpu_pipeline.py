"""Multi-step calls: drb_sim_step(sim, n) queues each step behind the previous one (the next layout
is read back mid-step).  The kernels and their order are those of n single-step calls, so the
results must be the same bit for bit; the point-update counter and the stage timers must add up."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle import v2
from oracle.physics import equilibrium

pytestmark = pytest.mark.gpu


def _obstacle_field(n=32, h=0.5, circle=(8.0, 8.0, 2.5)):
    """Jittered points around a circle, flowing at it (as the obstacle tests)."""
    rng = np.random.default_rng(4)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = (np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))).astype(np.float32)
    pos = pos[~v2.inside32(pos, circle)]
    u = np.zeros((len(pos), 2))
    u[:, 0] = 0.05
    f = np.stack([equilibrium(h * h, v) for v in u]).astype(np.float32)
    return pos, f, dict(tau=0.6, dx=4 * h, alpha=4.0, rho_rest=h * h, circle=circle)


def _cases():
    g = workloads.gaussian_pulse_two_level(n=64, amp=0.05, sigma_cells=4.0)
    yield "open_two_level", g.pos, g.f, dict(g.kw)
    yield ("open_obstacle",) + _obstacle_field()
    c = workloads.taylor_green(n=32, tau=0.7)
    yield "periodic", c.pos, c.f, dict(c.kw)


@pytest.mark.parametrize("case", list(_cases()), ids=lambda c: c[0])
def test_multi_step_call_equals_single_steps(case):
    name, pos, f, kw = case
    a = drb.GpuSimulator(pos, f, **kw)
    b = drb.GpuSimulator(pos, f, **kw)
    ua, ub = a.updates, b.updates
    counts = []
    a.step(6)
    for _ in range(6):
        counts.append(b.count())
        b.step()
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb), name
    da, db = a.diagnostics(), b.diagnostics()
    assert da == db
    assert a.updates - ua == b.updates - ub == sum(counts)


def test_batch_multi_step_call_equals_single_steps():
    cases = [workloads.gaussian_pulse_two_level(n=40, amp=0.05, sigma_cells=3.0),
             workloads.taylor_green(n=24, tau=0.7),
             workloads.gaussian_pulse_two_level(n=34, amp=0.05, sigma_cells=3.0, seed=2)]
    a = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    b = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    a.step(5)
    for _ in range(5):
        b.step()
    for (pa, fa), (pb, fb) in zip(a.moments(), b.moments()):
        assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def test_stage_timers_average_over_the_call():
    g = workloads.gaussian_pulse_two_level(n=64, amp=0.05, sigma_cells=4.0)
    sim = drb.GpuSimulator(g.pos, g.f, **g.kw)
    sim.set_timing(True)
    sim.step(4)
    ms = sim.stage_ms()
    assert all(v > 0.0 for v in ms.values()), ms
    sim.set_timing(True, stages=["gather"])
    sim.step(3)
    ms = sim.stage_ms()
    assert ms["gather"] > 0.0 and all(v == 0.0 for k, v in ms.items() if k != "gather"), ms
