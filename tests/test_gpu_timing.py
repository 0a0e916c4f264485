"""Stage timers (drb_sim_set_timing / drb_sim_stage_ms): every stage timed, or only a chosen set,
and timing does not change the step's results."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from direbm_amd._lib import DrbError

pytestmark = pytest.mark.gpu


def test_stage_timers_all_and_selected():
    c = workloads.gaussian_pulse_two_level(n=64, amp=0.05, sigma_cells=4.0)
    a = drb.GpuSimulator(c.pos, c.f, **c.kw)
    b = drb.GpuSimulator(c.pos, c.f, **c.kw)
    a.set_timing(True)
    a.step()
    ms = a.stage_ms()
    assert list(ms) == list(drb.GpuSimulator.STAGES)
    assert all(v > 0.0 for v in ms.values()), ms
    a.set_timing(True, stages=["gather", "bin"])
    a.step()
    ms = a.stage_ms()
    assert ms["gather"] > 0.0 and ms["bin"] > 0.0
    assert all(v == 0.0 for k, v in ms.items() if k not in ("gather", "bin")), ms
    a.set_timing(False)
    a.step()
    b.step(3)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)


def test_stage_timer_mask_is_checked():
    c = workloads.taylor_green(n=16)
    sim = drb.GpuSimulator(c.pos, c.f, **c.kw)
    for bad in (-1, 2, 0x100 | 1, 0x400 | 0x200 | 1):
        with pytest.raises(DrbError, match="INVALID_ARG"):
            drb._lib.call("drb_sim_set_timing", sim._h, bad)
