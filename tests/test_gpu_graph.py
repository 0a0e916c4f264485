"""Graph mode (captured HIP graphs of the periodic step): bit-identical to plain steps, replayed
several times with eager work and host reads between the replays; refused for open domains."""

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


def test_graph_refused_for_open_domain():
    pos, f = workloads.rest_pulse(4)
    sim = drb.GpuSimulator(pos, f, tau=0.6)
    with pytest.raises(drb.DrbError):
        sim.set_graph(True)
