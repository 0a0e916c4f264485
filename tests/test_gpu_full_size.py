"""Size-independent properties at the benchmark sizes, where the FP64 oracle would take too long:
determinism (two engines, identical bits), the mass budget m' = m - dropped + injected (FP64
reductions on the GPU), batched instances bit-identical to their own single runs, and the 3D step
on the 8M-point periodic field conserving mass."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads

pytestmark = pytest.mark.gpu


def test_full_size_two_level_pulse_deterministic_with_mass_budget():
    case = workloads.gaussian_pulse_two_level(n=1472)  # 1,001,296 points, the benchmark workload
    a = drb.GpuSimulator(case.pos, case.f, **case.kw)
    b = drb.GpuSimulator(case.pos, case.f, **case.kw)
    m_prev = a.diagnostics()["mass"]
    for _ in range(2):
        a.step()
        b.step()
        d = a.diagnostics()
        assert abs(d["mass"] - (m_prev - d["dropped"] + d["injected"])) <= 1e-6 * m_prev
        m_prev = d["mass"]
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert len(pa) > 1_000_000
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def test_large_batch_instances_match_single_runs_bitwise():
    cases = workloads.batch_cases(ninst=12, nmin=65536, nmax=262144, seed=11)
    batch = drb.BatchSimulator([(c.pos, c.f, c.kw) for c in cases])
    batch.step(2)
    got = batch.moments()
    for i, c in enumerate(cases):
        s = drb.GpuSimulator(c.pos, c.f, **c.kw)
        s.step(2)
        p, f = s.moments()
        assert np.array_equal(got[i][0], p) and np.array_equal(got[i][1], f), f"instance {i}"
        s.close()


def test_full_size_3d_periodic_field_deterministic_and_conserving():
    case = workloads.shear_layer_3d(n=200)  # 8,000,000 points
    a = drb.GpuSimulator3D(case.pos, case.f, **case.kw)
    m0 = a.diagnostics()["mass"]
    a.step(2)
    assert abs(a.diagnostics()["mass"] - m0) <= 1e-6 * m0
    b = drb.GpuSimulator3D(case.pos, case.f, **case.kw)
    b.step(2)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)
