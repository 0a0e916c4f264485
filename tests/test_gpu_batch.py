"""Batched instances (one launch per stage over a segment table) vs separate single runs.

Per-instance results must be bit-for-bit those of the instance's own GpuSimulator: the keys carry
each instance's cell offset and every kernel's arithmetic for an instance is independent of the
others.
"""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads

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


def test_batch_matches_single_runs_bitwise():
    cases = _mixed_cases()
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    singles = [drb.GpuSimulator(c.pos, c.f, **c.kw) for c in cases]
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
