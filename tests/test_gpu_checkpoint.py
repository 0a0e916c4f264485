"""Resume from a checkpoint: save after k steps, load, continue -- bit-identical to an
uninterrupted run (single simulation and a batch)."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads

pytestmark = pytest.mark.gpu


def test_resume_single_bit_identical(tmp_path):
    c = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(c.pos, c.f, **c.kw)
    a.step(5)
    b = drb.GpuSimulator(c.pos, c.f, **c.kw)
    b.step(3)
    b.save(tmp_path / "s.npz")
    b.close()
    r = drb.GpuSimulator.load(tmp_path / "s.npz")
    assert r.iteration == 3
    r.step(2)
    pa, fa = a.moments()
    pr, fr = r.moments()
    assert np.array_equal(pa, pr) and np.array_equal(fa, fr)


def test_resume_batch_bit_identical(tmp_path):
    cases = [(c.pos, c.f, c.kw) for c in (workloads.taylor_green(n=24, tau=0.7), workloads.shear_layer(n=20))]
    a = drb.BatchSimulator(cases)
    a.step(4)
    b = drb.BatchSimulator(cases)
    b.step(2)
    b.save(tmp_path / "b.npz")
    r = drb.BatchSimulator.load(tmp_path / "b.npz")
    r.step(2)
    for (pa, fa), (pr, fr) in zip(a.moments(), r.moments()):
        assert np.array_equal(pa, pr) and np.array_equal(fa, fr)


def test_resume_tuned_run_bit_identical(tmp_path):
    """A checkpoint keeps the tiles (lanes per point and candidates in flight: they set the order of
    the fp32 partial sums), so a run with non-default tiles resumes bit for bit."""
    c = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(c.pos, c.f, **c.kw)
    a.set_tiles(2, 2, 8, 4)
    a.step(5)
    b = drb.GpuSimulator(c.pos, c.f, **c.kw)
    b.set_tiles(2, 2, 8, 4)
    b.step(3)
    b.save(tmp_path / "t.npz")
    b.close()
    r = drb.GpuSimulator.load(tmp_path / "t.npz")
    assert r.tiles() == {"refine_lanes": 2, "refine_depth": 2, "gather_lanes": 8, "gather_depth": 4}
    r.step(2)
    pa, fa = a.moments()
    pr, fr = r.moments()
    assert np.array_equal(pa, pr) and np.array_equal(fa, fr)
