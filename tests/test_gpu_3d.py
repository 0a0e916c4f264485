"""3D D3Q13 engine against the FP64 3D oracle (oracle/v3.py): teacher-forced steps (same control
points, f relative L2 <= 1e-5, positions to fp32 rounding), open and periodic; determinism; FP64
mass conservation on a saturated periodic field; the reference's 3D behaviour tests."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle import v3
from oracle.constants import C3, EPS_EMPTY, W3
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def _prm(kw):
    return v3.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], r_res=kw.get("r_res"),
                     box=kw.get("box"), sphere=kw.get("sphere"), eps=kw.get("eps", EPS_EMPTY))


def test_d3q13_constants_are_the_reference_set_in_fp32():
    from direbm_amd.sim3d import d3q13
    c, w = d3q13()
    assert np.array_equal(c, C3.astype(np.float32)) and np.array_equal(w, W3.astype(np.float32))


def _teacher_forced(case, steps, lanes=None):
    kw = case.kw
    sim = drb.GpuSimulator3D(case.pos, case.f, **kw)
    if lanes is not None:
        sim.set_lanes(*lanes)
    prm = _prm(kw)
    L = kw.get("box")
    worst = 0.0
    for _ in range(steps):
        p0, f0 = sim.moments()
        ref = v3.step(p0, f0, prm)
        sim.step()
        p1, f1 = sim.moments()
        assert len(p1) == len(ref.pos), "control-point count differs"
        d = p1.astype(np.float64) - ref.pos
        if L is not None:
            d = np.where(d > L / 2, d - L, np.where(d < -L / 2, d + L, d))
        assert np.abs(d).max() <= 1e-5 * max(1.0, np.abs(ref.pos).max())
        e = rel_l2(f1, ref.f)
        worst = max(worst, e)
        assert e <= 1e-5, e
    return sim, worst


def test_teacher_forced_open_pulse():
    sim, worst = _teacher_forced(workloads.pulse_3d(n=10), 2)
    print(f"3D open pulse worst rel L2 {worst:.2e}")


def test_teacher_forced_periodic_shear():
    sim, worst = _teacher_forced(workloads.shear_layer_3d(n=12), 2)
    print(f"3D periodic shear worst rel L2 {worst:.2e}")


def test_teacher_forced_periodic_short_windows():
    # windows spanning under a quarter of the box: each range's periodic image shift is known from
    # its cell (the engine's fast minimum image), here with the obstacle as well
    case = workloads.shear_layer_3d(n=18, alpha=2.0, sphere=0.05)
    sim, worst = _teacher_forced(case, 1)
    print(f"3D periodic shear, short windows, worst rel L2 {worst:.2e}")


def test_deterministic_and_mass_conserving():
    c = workloads.shear_layer_3d(n=12)
    a = drb.GpuSimulator3D(c.pos, c.f, **c.kw)
    b = drb.GpuSimulator3D(c.pos, c.f, **c.kw)
    m0 = a.diagnostics()["mass"]
    a.step(3)
    b.step(3)
    pa, fa = a.moments()
    pb, fb = b.moments()
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb)
    assert abs(a.diagnostics()["mass"] - m0) <= 1e-6 * m0


def test_reference_3d_behaviour():
    # tests/test_reference_3d.py:9-40: a single pulse spreads; a rest block keeps its density
    f = v3.equilibrium(np.float64(1.5), np.zeros(3))[None, :]
    sim = drb.GpuSimulator3D(np.zeros((1, 3)), f, tau=0.6)
    sim.step(3)
    x, fx = sim.moments()
    assert len(x) > 1 and np.isfinite(fx).all() and (fx.sum(1) > 0).all()
    assert np.linalg.norm(x.astype(np.float64), axis=1).max() > 1.0
    g = np.arange(-2, 3, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    f = np.tile(v3.equilibrium(np.float64(1.0), np.zeros(3)), (len(pos), 1))
    sim = drb.GpuSimulator3D(pos, f, tau=0.6, rho_rest=1.0)
    sim.step()
    x, fx = sim.moments()
    interior = (np.abs(x) < 1.5).all(axis=1)
    assert abs(fx[interior].astype(np.float64).sum() / 27.0 - 1.0) < 0.3


def test_periodic_rejects_positions_outside_box():
    from direbm_amd._lib import DrbError
    c = workloads.shear_layer_3d(n=12)
    pos = c.pos.copy()
    pos[5, 2] = -0.01
    with pytest.raises(DrbError, match="periodic positions"):
        drb.GpuSimulator3D(pos, c.f, **c.kw)
    c2 = workloads.shear_layer(n=16)  # the 2D engine applies the same check
    p2 = c2.pos.copy()
    p2[3, 0] = c2.kw["box"] * 1.5
    with pytest.raises(DrbError, match="periodic positions"):
        drb.GpuSimulator(p2, c2.f, **c2.kw)


# ----------------------------------------------------------------------------- spherical obstacle
def _sphere_block():
    # tests/test_boundary_3d.py:29-46: a 5^3 block outside a unit sphere, one pulse point
    sph = (0.0, 0.0, 0.0, 1.0)
    f_rest = v3.equilibrium(np.float64(1.0), np.zeros(3))
    pulse = v3.equilibrium(np.float64(1.6), np.zeros(3))
    pos, f = [], []
    for i in range(-2, 3):
        for j in range(-2, 3):
            for k in range(-2, 3):
                x = np.array([float(i), float(j), float(k)])
                if v3.inside32(x[None], sph)[0]:
                    continue
                pos.append(x)
                f.append(pulse if (i, j, k) == (2, 0, 0) else f_rest)
    return np.array(pos, np.float32), np.array(f, np.float32), sph


def test_sphere_keeps_fluid_out_and_matches_the_oracle():
    pos, f, sph = _sphere_block()
    case = workloads.Case("sphere_block", pos, f, dict(tau=0.6, dx=1.0, alpha=4.0, rho_rest=1.0, sphere=sph))
    sim, worst = _teacher_forced(case, 2)
    x, fx = sim.moments()
    assert len(x) > 0 and not v3.inside32(x, sph).any() and np.isfinite(fx).all()
    print(f"sphere block worst rel L2 {worst:.2e}")


def test_sphere_in_periodic_shear_matches_the_oracle():
    case = workloads.shear_layer_3d(n=18, sphere=0.05)
    sim, worst = _teacher_forced(case, 1)
    print(f"periodic shear + sphere worst rel L2 {worst:.2e}")


def test_sphere_must_sit_inside_the_periodic_box():
    from direbm_amd._lib import DrbError
    c = workloads.shear_layer_3d(n=12)
    with pytest.raises(DrbError, match="obstacle"):
        drb.GpuSimulator3D(c.pos, c.f, sphere=(0.5, 0.5, 0.5, 0.1), **c.kw)


@pytest.mark.parametrize("lanes", [(8, 16), (16, 8)])
@pytest.mark.parametrize("which", ["open_pulse", "periodic_shear", "short_windows", "sphere_shear"])
def test_teacher_forced_lane_settings(which, lanes):
    """Every lane setting of the refine and gather kernels holds the oracle's per-step bar."""
    case = {"open_pulse": lambda: workloads.pulse_3d(n=10), "periodic_shear": lambda: workloads.shear_layer_3d(n=12),
            "short_windows": lambda: workloads.shear_layer_3d(n=18, alpha=2.0, sphere=0.05),
            "sphere_shear": lambda: workloads.shear_layer_3d(n=24, sphere=0.1)}[which]()
    _, worst = _teacher_forced(case, 2, lanes)
    print(f"3D {which} lanes {lanes}: worst rel L2 {worst:.2e}")
