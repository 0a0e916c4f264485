"""Autotuner of the refine / gather tiles (lanes per point, candidates in flight per lane): valid
choices, reproducible with set_tiles, and every tile keeps oracle parity (teacher-forced step,
relative L2 <= 1e-5)."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle import v2
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def test_autotune_reproducible_with_set_tiles():
    c = workloads.taylor_green(n=48, tau=0.8)
    a = drb.GpuSimulator(c.pos, c.f, **c.kw)
    tiles = a.autotune()
    assert tiles["refine_lanes"] in (2, 4, 8) and tiles["gather_lanes"] in (2, 4, 8)
    assert tiles["refine_depth"] in (2, 4) and tiles["gather_depth"] in (2, 4)
    assert a.tiles() == tiles
    a.step(3)
    b = drb.GpuSimulator(c.pos, c.f, **c.kw)
    b.step(1)  # the tuning step runs with the default tiles
    b.set_tiles(**tiles)
    b.step(3)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


@pytest.mark.parametrize("tiles", [(2, 4, 2, 4), (8, 4, 8, 4), (2, 4, 8, 4), (4, 2, 4, 2), (8, 2, 2, 4)])
def test_tile_settings_keep_parity(tiles):
    c = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    sim = drb.GpuSimulator(c.pos, c.f, **c.kw)
    sim.set_tiles(*tiles)
    kw = c.kw
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], two_level=kw["two_level"])
    for _ in range(2):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos)
        assert rel_l2(f1, ref.f) <= 1e-5
        assert np.abs(p1 - ref.pos).max() <= 1e-5 * max(1.0, np.abs(ref.pos).max())


def test_autotune_leaves_the_step_diagnostics():
    """The tuning relaunches overwrite the step's buffers; the tuner restores them, so the dropped
    and injected mass of the tuning step are those of an untuned step (same default tiles)."""
    c = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(c.pos, c.f, **c.kw)
    a.autotune()
    b = drb.GpuSimulator(c.pos, c.f, **c.kw)
    b.step()
    da, db = a.diagnostics(), b.diagnostics()
    assert db["injected"] > 0.0
    for key in ("mass", "mom_x", "mom_y", "dropped", "injected"):
        assert da[key] == db[key], key
