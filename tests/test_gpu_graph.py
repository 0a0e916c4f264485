"""Graph mode (captured HIP graphs of the step): bit-identical to plain steps, replayed several
times with eager work and host reads between the replays -- periodic boxes, open domains laid out
on the device (one-level and two-level), mixed batches, and open clouds that outgrow the graph's
cell capacity mid-call (the call resumes from the step that ran empty)."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads

pytestmark = pytest.mark.gpu


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_graph_single_bit_identical():
    c = workloads.taylor_green(n=40, tau=0.8)
    plain = drb.GpuSimulator(c.pos, c.f, **c.kw)
    graph = drb.GpuSimulator(c.pos, c.f, **c.kw)
    graph.set_graph(True)
    for _ in range(4):  # one replay at a time, with host reads between
        plain.step()
        graph.step()
        assert plain.count() == graph.count()
        assert _same(plain.moments(), graph.moments())
    plain.step(5)
    graph.step(5)
    assert _same(plain.moments(), graph.moments())
    assert plain.diagnostics()["mass"] == graph.diagnostics()["mass"]


def test_graph_batch_bit_identical():
    cases = [(c.pos, c.f, c.kw) for c in (workloads.taylor_green(n=24, tau=0.7), workloads.shear_layer(n=30),
                                          workloads.taylor_green(n=16, tau=1.2, seed=3))]
    plain = drb.BatchSimulator(cases)
    graph = drb.BatchSimulator(cases)
    graph.set_graph(True)
    for n in (1, 3, 2):
        plain.step(n)
        graph.step(n)
        assert list(plain.counts()) == list(graph.counts())
        for a, b in zip(plain.moments(), graph.moments()):
            assert _same(a, b)
    # replayed steps count their point-updates (the moments entering each step) as plain steps do
    assert graph.updates == plain.updates > 0


def _pair(case, horizon=None):
    plain = drb.GpuSimulator(case.pos, case.f, **case.kw)
    graph = drb.GpuSimulator(case.pos, case.f, **case.kw)
    graph.set_graph(True, horizon=horizon)
    return plain, graph


def test_graph_open_two_level_bit_identical():
    """The headline workload's form: an open, growing two-level cloud, each step's grid laid out
    on the device from the previous step's boxes."""
    plain, graph = _pair(workloads.gaussian_pulse_two_level(n=96, amp=0.05, sigma_cells=4.0))
    for n in (1, 3, 6):
        plain.step(n)
        graph.step(n)
        assert plain.count() == graph.count()
        assert _same(plain.moments(), graph.moments())
    dp, dg = plain.diagnostics(), graph.diagnostics()
    for key in ("mass", "mom_x", "mom_y", "dropped", "injected"):
        assert dp[key] == dg[key], key
    assert graph.updates == plain.updates > 0
    assert graph.graph_info()["captures"] >= 1
    plain.step(2)  # plain steps after replays read the boxes the replays left
    graph.set_graph(False)
    graph.step(2)
    assert _same(plain.moments(), graph.moments())


def test_graph_open_resumes_past_capacity():
    """Horizon 1: the capacity holds only the current grid, so a later step of the call finds its
    grown grid beyond it, runs empty, and the call resumes from that step with a new capture; the
    result equals plain steps bit for bit."""
    c = workloads.gaussian_pulse_two_level(n=64, amp=0.05, sigma_cells=4.0)
    plain, graph = _pair(c, horizon=1)
    plain.step(12)
    graph.step(12)
    info = graph.graph_info()
    assert info["resumes"] >= 1 and info["captures"] >= 2
    assert plain.count() == graph.count()
    assert _same(plain.moments(), graph.moments())
    assert graph.updates == plain.updates
    # one-level open grid, small enough that its ring of new points per step outgrows the point
    # capacity (the moments and a quarter more) as well as the cells
    pos, f = workloads.rest_pulse(6, pulse_rho=1.5)
    a = drb.GpuSimulator(pos, f, tau=0.6)
    b = drb.GpuSimulator(pos, f, tau=0.6)
    b.set_graph(True, horizon=1)
    for n in (5, 4, 20):
        a.step(n)
        b.step(n)
        assert a.count() == b.count()
        assert _same(a.moments(), b.moments())
    assert b.graph_info()["resumes"] >= 2
    assert a.updates == b.updates


def test_graph_open_pulse_outgrows_point_capacity():
    """Default horizon: a small open pulse grows from 12k to 33k points in 25 steps, past the
    control-point capacity (the moments and a sixteenth more) several times within one call; each
    time the step that found too many points ran empty and the call resumed from it."""
    plain, graph = _pair(workloads.gaussian_pulse_two_level(n=160))
    plain.step(25)
    graph.step(25)
    assert graph.graph_info()["resumes"] >= 1
    assert plain.count() == graph.count()
    assert _same(plain.moments(), graph.moments())
    assert graph.updates == plain.updates


def test_graph_mixed_batch_bit_identical():
    """Open and periodic instances in one batch: the periodic grids stay fixed, the open ones are
    laid out on the device, the cell offsets follow from both."""
    pulse = workloads.gaussian_pulse_two_level(n=48, amp=0.05, sigma_cells=4.0)
    pos, f = workloads.rest_pulse(5, pulse_rho=1.2)
    tg = workloads.taylor_green(n=24, tau=0.7)
    cases = [(pulse.pos, pulse.f, pulse.kw), (tg.pos, tg.f, tg.kw), (pos, f, {"tau": 0.6})]
    plain = drb.BatchSimulator(cases)
    graph = drb.BatchSimulator(cases)
    graph.set_graph(True)
    for n in (2, 5, 1):
        plain.step(n)
        graph.step(n)
        assert list(plain.counts()) == list(graph.counts())
        for a, b in zip(plain.moments(), graph.moments()):
            assert _same(a, b)
        # per-instance mass, momentum, dropped and injected mass of the last step: read with the
        # segment table that step ran with
        assert np.array_equal(plain.diagnostics(), graph.diagnostics())
    assert graph.updates == plain.updates > 0


def test_graph_open_instance_without_moments():
    """An open instance that holds no moments keeps an empty grid in the device layout."""
    pulse = workloads.gaussian_pulse_two_level(n=32, amp=0.05, sigma_cells=4.0)
    empty = (np.zeros((0, 2), np.float32), np.zeros((0, 7), np.float32), {"tau": 0.7})
    cases = [empty, (pulse.pos, pulse.f, pulse.kw)]
    plain = drb.BatchSimulator(cases)
    graph = drb.BatchSimulator(cases)
    graph.set_graph(True)
    plain.step(3)
    graph.step(3)
    assert list(graph.counts()) == list(plain.counts()) and graph.counts()[0] == 0
    for a, b in zip(plain.moments(), graph.moments()):
        assert _same(a, b)
