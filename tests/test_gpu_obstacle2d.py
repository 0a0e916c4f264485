"""The 2D circular obstacle on the GPU (drb_sim_set_obstacle): teacher-forced steps against the
FP64 oracle (open and periodic, with a density pulse next to the circle so that bounced parts
carry different values), the reference's own keep-fluid-out scenario, a batched instance with an
obstacle bit-identical to its single run, and the modes that refuse obstacles."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd._lib import DrbError
from oracle import v2
from oracle.physics import equilibrium
from tests._util import min_image, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _field(periodic, n=32, h=0.5, circle=(8.0, 8.0, 2.5), seed=0):
    """Jittered n x n points (spacing h) around a circle, a Gaussian density bump just left of it."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = (np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))).astype(np.float32)
    pos = pos[~v2.inside32(pos, circle)]
    r2 = ((pos - np.array([4.5, 8.0])) ** 2).sum(1)
    rho = h * h * (1.0 + 0.2 * np.exp(-r2 / 2.0))
    u = np.zeros((len(pos), 2))
    u[:, 0] = 0.05
    f = np.stack([equilibrium(r, v) for r, v in zip(rho, u)]).astype(np.float32)
    kw = dict(tau=0.6, dx=4 * h, alpha=4.0, rho_rest=h * h, circle=circle)
    if periodic:
        kw["box"] = n * h
    return pos, f, kw


def _oracle(kw):
    return v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], box=kw.get("box"),
                     circle=kw["circle"])


@pytest.mark.parametrize("periodic", [False, True])
def test_teacher_forced_steps_with_a_circle(periodic):
    pos, f, kw = _field(periodic)
    sim = drb.GpuSimulator(pos, f, **kw)
    prm = _oracle(kw)
    for _ in range(3):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos), "moment count differs"
        d = np.abs(min_image(p1.astype(np.float64) - ref.pos, kw.get("box")))
        assert d.max() <= 1e-5 * max(1.0, np.abs(ref.pos).max())
        e = rel_l2(f1, ref.f)
        assert e <= TOL, f"relative L2 {e:.3g}"
        assert not v2.inside32(p1, kw["circle"]).any()
    # the scenario does bounce: the oracle's component count exceeds 7 per moment
    p0, f0 = sim.moments()
    pc, _, _ = v2.disperse(p0, v2.collide(f0.astype(np.float64), kw["tau"]), kw["dx"], prm)
    assert len(pc) > 7 * len(p0)


def test_reference_keep_fluid_out_scenario():
    # the reference's tests/test_boundary.py:56-72 (there on its v1 solver), here on the engine
    circle = (0.0, 0.0, 2.0)
    xs, fs = [], []
    for i in range(-6, 7):
        for j in range(-6, 7):
            x = np.array([float(i), float(j)])
            if v2.inside32(x[None], circle)[0]:
                continue
            rho = 1.5 if (i == 4 and j == 0) else 1.0
            xs.append(x)
            fs.append(equilibrium(np.float64(rho), np.zeros(2)))
    sim = drb.GpuSimulator(np.array(xs), np.array(fs), tau=0.6, rho_rest=1.0, circle=circle)
    for _ in range(5):
        sim.step()
    pos, f = sim.moments()
    assert len(pos) > 0
    assert not v2.inside32(pos, circle).any()
    assert np.isfinite(f).all()


def test_batched_instance_with_circle_matches_its_single_run():
    a = _field(False)
    b = _field(True, seed=1)
    c = _field(False, circle=(6.0, 9.0, 1.5), seed=2)
    plain = (a[0], a[1], {k: v for k, v in a[2].items() if k != "circle"})
    batch = drb.BatchSimulator([a, b, plain, c])
    batch.step(2)
    got = batch.moments()
    for i, case in enumerate((a, b, plain, c)):
        s = drb.GpuSimulator(case[0], case[1], **case[2])
        s.step(2)
        p, f = s.moments()
        assert np.array_equal(got[i][0], p) and np.array_equal(got[i][1], f), f"instance {i}"


def test_obstacle_refusals():
    pos, f, kw = _field(True)
    sim = drb.GpuSimulator(pos, f, **kw)
    with pytest.raises(DrbError, match="INVALID_ARG"):
        sim.set_graph(True)
    with pytest.raises(DrbError, match="INVALID_ARG"):  # a periodic circle must keep dx off the edges
        drb.GpuSimulator(pos, f, **dict(kw, circle=(0.5, 8.0, 1.0)))
    with pytest.raises(ValueError):
        drb.GpuSimulator(pos, f, **dict(kw, circle=(8.0, 8.0, 0.0)))


def test_teacher_forced_two_level_with_a_circle():
    # two-level thinning (fine cells inside r_in of the centre, coarse outside) and a resample
    # radius of 1.5 cells, around a circle that straddles the level boundary
    pos, f, kw = _field(False, circle=(9.0, 8.0, 2.0))
    kw = dict(kw, two_level=(8.0, 8.0, 4.0), r_res=1.5)
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], circle=kw["circle"],
                    two_level=kw["two_level"], r_res=kw["r_res"])
    sim = drb.GpuSimulator(pos, f, **kw)
    for _ in range(2):
        p0, f0 = sim.moments()
        ref = v2.step(p0, f0, prm)
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos), "moment count differs"
        assert np.abs(p1.astype(np.float64) - ref.pos).max() <= 1e-5 * max(1.0, np.abs(ref.pos).max())
        assert rel_l2(f1, ref.f) <= TOL
        assert not v2.inside32(p1, kw["circle"]).any()


def test_points_inside_a_circle_are_removed_and_the_empty_state_steps_on():
    # every point and every component lies inside the circle: the moving components find no
    # crossing and are lost (as in the reference), the rest components form points inside, which
    # are removed; later steps are no-ops
    g = np.linspace(7.0, 9.0, 5)
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    f = np.tile(equilibrium(np.float64(0.25), np.zeros(2)), (len(pos), 1)).astype(np.float32)
    sim = drb.GpuSimulator(pos, f, tau=0.6, dx=2.0, rho_rest=0.25, circle=(8.0, 8.0, 4.0))
    sim.step()
    assert sim.count() == 0
    sim.step(2)
    p, ff = sim.moments()
    assert p.shape == (0, 2) and ff.shape == (0, 7)
