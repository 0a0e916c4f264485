"""Batched instances (one launch per stage over a segment table) vs separate single runs and vs
the FP64 oracle.

Per-instance results must be bit-for-bit those of the instance's own GpuSimulator: the keys carry
each instance's cell offset and every kernel's arithmetic for an instance is independent of the
others.  Each instance must also hold the oracle's per-step parity bar (same control points,
f relative L2 <= 1e-5), since the reference steps one cloud at a time.
"""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle import v2
from tests._util import min_image, rel_l2

pytestmark = pytest.mark.gpu


def _mixed_cases():
    cases = [workloads.taylor_green(n=24, tau=0.7), workloads.shear_layer(n=40, tau=1.2, seed=3),
             workloads.taylor_green(n=33, u0=0.08, tau=0.55, seed=5)]
    pos, f = workloads.rest_pulse(7, pulse_rho=1.5)
    cases.append(workloads.Case("rest_pulse", pos.astype(np.float32), f.astype(np.float32),
                                dict(tau=0.6, rho_rest=1.0)))
    g = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    cases.append(g)
    return cases


# lanes (8, 8): groups of eight lanes per point, with an open two-level instance added; the batch
# and its single runs share the lanes (they set the order of the fp32 partial sums).
@pytest.mark.parametrize("lanes", [None, (8, 8)])
def test_batch_matches_single_runs_bitwise(lanes):
    cases = _mixed_cases()
    if lanes is not None:
        cases.append(workloads.gaussian_pulse_two_level(n=34, amp=0.05, sigma_cells=3.0, seed=2))
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    singles = [drb.GpuSimulator(c.pos, c.f, **c.kw) for c in cases]
    if lanes is not None:
        batch.set_lanes(*lanes)
        for s in singles:
            s.set_lanes(*lanes)
    for _ in range(3):
        batch.step()
        for s in singles:
            s.step()
        got = batch.moments()
        for b, s in enumerate(singles):
            p, f = s.moments()
            assert np.array_equal(got[b][0], p), f"instance {b} positions"
            assert np.array_equal(got[b][1], f), f"instance {b} distributions"
    d = batch.diagnostics()
    for b, s in enumerate(singles):
        assert np.isclose(d[b, 0], s.diagnostics()["mass"], rtol=1e-12)


def test_batch_of_generated_instances_conserves_mass():
    cases = workloads.batch_cases(ninst=12, nmin=256, nmax=4096, seed=7)
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    m0 = batch.diagnostics()[:, 0]
    n0 = batch.counts()
    batch.step(4)
    d = batch.diagnostics()
    assert np.array_equal(batch.counts(), n0)  # saturated periodic boxes keep one point per cell
    assert np.allclose(d[:, 0], m0, rtol=1e-6)
    assert (d[:, 3] == 0).all()


def _oracle_params(kw):
    return v2.Params(tau=kw["tau"], dx=kw.get("dx", 1.0), alpha=kw.get("alpha", 4.0),
                     kappa_hard=kw.get("kappa_hard", 4), rho_rest=kw.get("rho_rest", 1.0),
                     eps=kw.get("eps") or v2.EPS_EMPTY, r_res=kw.get("r_res"), box=kw.get("box"),
                     two_level=kw.get("two_level"), circle=kw.get("circle"))


def _check_against_oracle(p1, f1, ref, kw, what):
    assert len(p1) == len(ref.pos), f"{what}: moment count"
    d = np.abs(min_image(p1.astype(np.float64) - ref.pos, kw.get("box")))
    assert d.max() <= 1e-5 * max(1.0, np.abs(ref.pos).max()), f"{what}: positions"
    e = rel_l2(f1, ref.f)
    assert e <= 1e-5, f"{what}: relative L2 {e:.3g}"


def test_batch_instances_teacher_forced_against_oracle():
    """Every instance of a mixed batch (periodic Taylor-Green, double shear layer, resample radius
    1.5 h, open two-level pulse, circles in an open and a periodic box) against the FP64 oracle's
    step from the batch's own state, three steps: the per-instance semantics of
    direbm/warp/simulator.py:50-73."""
    from tests.test_gpu_obstacle2d import _field

    tg = workloads.taylor_green(n=32, tau=0.8, seed=2)
    cases = [workloads.taylor_green(n=24, tau=0.7), workloads.shear_layer(n=40, tau=1.2, seed=3),
             workloads.taylor_green(n=32, tau=0.8, seed=2, r_res=1.5 * tg.kw["dx"] / 4.0),
             workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)]
    cases = [(c.pos, c.f, c.kw) for c in cases] + [_field(False), _field(True, seed=1)]
    batch = drb.BatchSimulator(cases)
    prms = [_oracle_params(kw) for _, _, kw in cases]
    for step in range(3):
        state = batch.moments()
        refs = [v2.step(p0, f0, prm) for (p0, f0), prm in zip(state, prms)]
        batch.step()
        got = batch.moments()
        diag = batch.diagnostics()
        for b, ((p1, f1), ref) in enumerate(zip(got, refs)):
            _check_against_oracle(p1, f1, ref, cases[b][2], f"step {step} instance {b}")
            m_ref = float(ref.f.sum())
            assert abs(diag[b, 0] - m_ref) <= 1e-6 * m_ref, f"step {step} instance {b}: FP64 mass"


def test_config4_batch_256_instances():
    """The 256-instance batch of the batched configuration (workloads.batch_cases(256): periodic
    Taylor-Green / double shear layers of 4k-256k points, 15.2M points in all), two steps: every
    instance bit-identical to its own single run, FP64 mass of every instance unchanged to 1e-6,
    and the three smallest instances teacher-forced against the FP64 oracle."""
    cases = workloads.batch_cases(ninst=256)
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    assert batch.count() > 15_000_000
    m0 = batch.diagnostics()[:, 0].copy()
    small = np.argsort([len(c.pos) for c in cases])[:3]
    prev = batch.moments()
    for step in range(2):
        refs = {int(b): v2.step(prev[b][0], prev[b][1], _oracle_params(cases[b].kw)) for b in small}
        batch.step()
        prev = batch.moments()
        for b, ref in refs.items():
            _check_against_oracle(prev[b][0], prev[b][1], ref, cases[b].kw, f"step {step} instance {b}")
    d = batch.diagnostics()
    assert np.all(np.abs(d[:, 0] - m0) <= 1e-6 * m0)
    for b, c in enumerate(cases):
        s = drb.GpuSimulator(c.pos, c.f, **c.kw)
        s.step(2)
        p, f = s.moments()
        assert np.array_equal(prev[b][0], p) and np.array_equal(prev[b][1], f), f"instance {b}"
        s.close()
