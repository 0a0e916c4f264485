"""The oracle's 2D obstacle (circle, specular bounce, adjacent-direction split) against the
reference's own known answers (tests/test_boundary.py of the reference: inside test, entry point
and outward normal, a miss, specular reflection, split weights) and the invariants that fix the
rest: a bounce conserves the component's value, its parts point away from the surface, and no
moment is left inside after a step."""

import numpy as np
import pytest

from oracle import v2
from oracle.constants import C2
from oracle.physics import equilibrium

UNIT = (0.0, 0.0, 1.0)


def test_inside_is_strict():
    assert v2.inside32(np.array([[0.0, 0.0]]), UNIT)[0]
    assert not v2.inside32(np.array([[2.0, 0.0]]), UNIT)[0]
    assert not v2.inside32(np.array([[1.0, 0.0]]), UNIT)[0]  # on the circle is outside


def test_ray_hit_entry_point_and_outward_normal():
    hit, n = v2.ray_hit64((2.0, 0.0), (0.0, 0.0), UNIT)
    assert hit and np.allclose(n, [1.0, 0.0])
    hit, n = v2.ray_hit64((0.0, -3.0), (0.0, 0.0), UNIT)
    assert hit and np.allclose(n, [0.0, -1.0])


def test_ray_miss():
    assert not v2.ray_hit64((2.0, 2.0), (2.0, -2.0), UNIT)[0]
    assert not v2.ray_hit64((3.0, 0.0), (2.0, 0.0), UNIT)[0]  # stops short


def test_reflect_specular():
    assert np.allclose(v2.reflect64((-1.0, 0.0), (1.0, 0.0)), [1.0, 0.0])
    assert np.allclose(v2.reflect64((1.0, -1.0), (0.0, 1.0)), [1.0, 1.0])
    d = v2.reflect64(C2[2], np.array([0.6, -0.8]))
    assert abs(np.hypot(*d) - 1.0) < 1e-15  # length kept


def test_split_weights_and_aligned_directions():
    for ang in np.linspace(0, 2 * np.pi, 17):
        parts = v2.split_hex(np.array([np.cos(ang), np.sin(ang)]))
        assert abs(sum(w for _, w in parts) - 1.0) < 1e-12
        assert all(1 <= j <= 6 for j, _ in parts)
    for j in range(1, 7):  # exactly along c_j: all weight on j
        parts = dict(v2.split_hex(C2[j]))
        assert abs(parts[j] - 1.0) < 1e-9
    parts = dict(v2.split_hex(np.array([np.cos(-np.pi / 6), np.sin(-np.pi / 6)])))  # -30 degrees
    assert abs(parts[6] - 0.5) < 1e-12 and abs(parts[1] - 0.5) < 1e-12


def test_bounce_conserves_value_and_points_away():
    # one moment left of a circle: its directions 1, 2 and 6 land inside and bounce
    prm = v2.Params(tau=0.6, dx=1.0, circle=(2.0, 0.0, 1.4))
    pos = np.array([[0.5, 0.0]], np.float32)
    f = np.arange(1, 8, dtype=np.float64)[None] / 10
    ends = (pos + C2.astype(np.float32)).astype(np.float32)
    hit = v2.inside32(ends, prm.circle)
    assert list(np.flatnonzero(hit)) == [1, 2, 6]
    pc, val, dirs = v2.disperse(pos, f, 1.0, prm)
    assert abs(val.sum() - f.sum()) < 1e-14  # every bounced component hits the circle here
    assert not v2.inside32(pc, prm.circle).any()
    assert list(dirs[:4]) == [0, 3, 4, 5]  # the regular components come first
    assert len(dirs) == 4 + 2 * 3  # two parts per bounce
    # each reflection points out of the circle, and its two parts straddle it
    for k in (1, 2, 6):
        ok, n = v2.ray_hit64(pos[0], ends[k], prm.circle)
        d = v2.reflect64(C2[k], n)
        assert ok and d @ n > 0
        mean = sum(w * C2[j] for j, w in v2.split_hex(d))
        assert mean @ d > np.cos(np.pi / 6) * np.hypot(*mean)  # within the 60-degree sector


def test_step_keeps_fluid_out():
    # the reference's own scenario (tests/test_boundary.py:56-72) on the oracle of the GPU step
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
    sim = v2.Simulator(np.array(xs), np.array(fs), tau=0.6, rho_rest=1.0, circle=circle)
    for _ in range(5):
        sim.step()
    pos, f = sim.moments()
    assert len(pos) > 0
    assert not v2.inside32(pos, circle).any()
    assert np.isfinite(f).all()


@pytest.mark.parametrize("periodic", [False, True])
def test_bounced_mass_reaches_the_control_points(periodic):
    # one step's scatter sees the parts: without gap-fill (rest state everywhere is not empty) the
    # scattered mass equals the dispersed mass minus what no control point reaches
    kw = dict(box=16.0) if periodic else {}
    g = (np.arange(32) + 0.5) * 0.5
    X, Y = np.meshgrid(g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel()], 1).astype(np.float32)
    circle = (8.0, 8.0, 2.5)
    pos = pos[~v2.inside32(pos, circle)]
    f = np.tile(equilibrium(np.float64(0.25), np.zeros(2)), (len(pos), 1))
    prm = v2.Params(tau=0.6, dx=1.0, rho_rest=0.25, circle=circle, **kw)
    r = v2.step(pos, f, prm)
    disp_mass = r.f_star.sum()
    assert abs(r.res.cp_f.sum() - (disp_mass - r.res.dropped)) < 1e-9 * disp_mass
    assert not v2.inside32(r.pos.astype(np.float32), circle).any()
