"""Edge cases of the GPU path: empty inputs at every stage and in the simulators, an empty
instance inside a batch, isolated control points (closed form), and the errors raised for
invalid parameters and non-finite state (forced from the host: every kernel still exits
normally, the engine reports the condition through its status word)."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from direbm_amd._lib import DrbError
from oracle.constants import W2 as W
from oracle.physics import equilibrium

pytestmark = pytest.mark.gpu

F32 = np.float32


def test_stages_on_empty_inputs():
    pos_c, val_c, dir_c = drb.disperse(np.zeros((0, 2), F32), np.zeros((0, 7), F32), 1.0)
    assert pos_c.shape == (0, 2) and val_c.shape == (0,) and dir_c.shape == (0,)
    cp, p = drb.create_control_points(np.zeros((0, 2), F32), 0.25)
    assert p == 0 and cp.shape == (0, 2)
    new = drb.refine_control_points(np.zeros((0, 2), F32), np.zeros((0, 2), F32), np.zeros(0, F32),
                                    np.zeros(0, np.int32), 1.0, 4)
    assert new.shape == (0, 2)
    out_pos, out_f = drb.resample(np.zeros((0, 2), F32), np.zeros(0, F32), np.zeros(0, np.int32),
                                  np.zeros((0, 2), F32), 1.0, W.astype(F32))
    assert out_pos.shape == (0, 2) and out_f.shape == (0, 7)


def test_refine_without_components_keeps_points():
    cp = np.array([[0.0, 0.0], [5.0, 1.0]], F32)
    new = drb.refine_control_points(cp, np.zeros((0, 2), F32), np.zeros(0, F32), np.zeros(0, np.int32), 1.0, 4)
    assert np.array_equal(new.numpy(), cp)  # kappa = 0 <= kappa_hard: kept


def test_isolated_control_points_get_the_rest_state():
    # no component within r: every direction is empty, and an isolated point's gap-fill sum is its
    # own weight r, so f = (r / r) f_rest = f_rest exactly (direbm/warp/propagation.py:203-218)
    cp = np.array([[0.0, 0.0], [10.0, 0.0], [0.0, 10.0]], F32)
    f_rest = (0.7 * W).astype(F32)
    out_pos, out_f = drb.resample(np.zeros((0, 2), F32), np.zeros(0, F32), np.zeros(0, np.int32), cp, 1.0, f_rest)
    assert np.array_equal(out_pos.numpy(), cp)
    assert np.array_equal(out_f.numpy(), np.tile(f_rest, (3, 1)))


def test_simulator_with_empty_state():
    sim = drb.GpuSimulator(np.zeros((0, 2), F32), np.zeros((0, 7), F32), tau=0.6)
    sim.step(3)
    pos, f = sim.moments()
    assert sim.count() == 0 and pos.shape == (0, 2) and f.shape == (0, 7)
    sim3 = drb.GpuSimulator3D(np.zeros((0, 3), F32), np.zeros((0, 13), F32), tau=0.6)
    sim3.step(2)
    assert sim3.count() == 0


def test_batch_with_an_empty_instance():
    pos, f = workloads.rest_pulse(5, pulse_rho=1.4)
    live = (pos.astype(F32), f.astype(F32), dict(tau=0.6, rho_rest=1.0))
    empty = (np.zeros((0, 2), F32), np.zeros((0, 7), F32), dict(tau=0.6, rho_rest=1.0))
    batch = drb.BatchSimulator([live, empty, live])
    single = drb.GpuSimulator(live[0], live[1], **live[2])
    batch.step(3)
    single.step(3)
    got = batch.moments()
    p1, f1 = single.moments()
    assert batch.counts()[1] == 0 and got[1][0].shape == (0, 2)
    for b in (0, 2):
        assert np.array_equal(got[b][0], p1) and np.array_equal(got[b][1], f1)


def test_invalid_parameters_raise():
    pos, f = np.zeros((1, 2), F32), equilibrium(np.float64(1.0), np.zeros(2))[None].astype(F32)
    for kw in (dict(tau=0.0), dict(tau=0.6, dx=0.0), dict(tau=0.6, alpha=-1.0)):
        with pytest.raises(DrbError, match="INVALID_ARG"):
            drb.GpuSimulator(pos, f, **kw)
    with pytest.raises(DrbError, match="INVALID_ARG"):
        drb.GpuSimulator3D(np.zeros((1, 3), F32), np.zeros((1, 13), F32), tau=-1.0)
    with pytest.raises(DrbError, match="INVALID_ARG"):
        drb.GpuSimulator(pos, f, tau=0.6, box=5.3)  # not a whole number of thinning cells


def test_stage_directions_outside_the_lattice_raise():
    # a component direction indexes one of the seven sums: 7 or -1 is rejected before any kernel
    # uses it
    pos_c = np.array([[0.0, 0.0], [0.5, 0.0]], F32)
    val_c = np.array([0.1, 0.2], F32)
    cp = np.array([[0.2, 0.0]], F32)
    for bad in (7, -1):
        dir_c = np.array([0, bad], np.int32)
        with pytest.raises(DrbError, match="INVALID_ARG"):
            drb.resample(pos_c, val_c, dir_c, cp, 1.0, W.astype(F32))
        with pytest.raises(DrbError, match="INVALID_ARG"):
            drb.refine_control_points(cp, pos_c, val_c, dir_c, 1.0, 4)
    out_pos, out_f = drb.resample(pos_c, val_c, np.array([0, 6], np.int32), cp, 1.0, W.astype(F32))
    assert out_f.shape == (1, 7)


def test_overflowing_refine_weights_are_reported():
    # populations near 500 after collide make refine's weights exp(val) overflow to inf, so the
    # weighted means are NaN; the step reports it (and bounds every window while doing so), and
    # the state stays the one before the step
    pos, f = workloads.rest_pulse(8)
    f = f.astype(F32)
    f[len(f) // 2, :] = 100.0
    sim = drb.GpuSimulator(pos.astype(F32), f, tau=0.6)
    m = sim.count()
    with pytest.raises(DrbError, match="NONFINITE"):
        sim.step()
    assert sim.count() == m
    c = workloads.taylor_green(n=32)
    f = c.f.astype(F32)
    f[100, :] = 100.0
    sim = drb.GpuSimulator(c.pos, f, **c.kw)
    with pytest.raises(DrbError, match="NONFINITE"):
        sim.step()


def test_nonfinite_state_is_reported():
    pos, f = workloads.rest_pulse(3)
    bad = pos.astype(F32)
    bad[4, 1] = np.nan
    with pytest.raises(DrbError, match="NONFINITE"):
        drb.GpuSimulator(bad, f.astype(F32), tau=0.6)
    fb = f.astype(F32)
    fb[2, 3] = np.inf
    sim = drb.GpuSimulator(pos.astype(F32), fb, tau=0.6)
    with pytest.raises(DrbError, match="NONFINITE"):
        sim.step()
