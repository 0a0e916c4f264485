"""Pins of the GPU-step oracle (oracle/v2.py) on the CPU.

Each check is independent of the oracle's own formulas: the reference's known answers
(tests/test_warp_propagation.py), pure-Python brute force over all pairs, and invariants the
algorithm fixes (mass balance of the scatter, isolated-point gap-fill, exact conservation on a
saturated periodic field, rest-field preservation).
"""

import numpy as np
import pytest

from direbm_amd import workloads
from oracle import v2
from oracle.constants import C2, W2
from oracle.fields import bin_fields
from oracle.physics import equilibrium
from tests import _brute
from tests._util import random_cloud


def _ring(center, radius, dirs, vals):
    ang = np.linspace(0.0, 2 * np.pi, len(dirs), endpoint=False)
    pos = np.array(center) + radius * np.c_[np.cos(ang), np.sin(ang)]
    return pos, np.array(dirs, dtype=np.int32), np.array(vals, dtype=np.float64)


def test_disperse_closed_form():
    pos_m, f_m = random_cloud(20, 0)
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    exp = (pos_m[:, None, :] + C2[None, :, :]).reshape(-1, 2)
    assert np.allclose(pc, exp, atol=1e-5)
    assert np.array_equal(vc, f_m.reshape(-1))
    assert np.array_equal(dc, np.tile(np.arange(7), 20))
    # component 7i+k is exactly fp32(pos + fp32(C_k * dx))
    i, k = 5, 3
    x = np.float32(pos_m[i, 0].astype(np.float32) + np.float32(np.float32(C2[k, 0]) * np.float32(1.0)))
    assert pc[7 * i + k, 0] == x


def test_thin_matches_reference_unique_restatement():
    # tests/test_warp_propagation.py:27-36,55-67
    pos_m, f_m = random_cloud(50, 1)
    pc, _, _ = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    cells = np.floor(pc / 0.25).astype(np.int64)
    keys = cells[:, 0] * 1_000_003 + cells[:, 1]
    uniq, inv = np.unique(keys, return_inverse=True)
    sums = np.zeros((len(uniq), 2))
    np.add.at(sums, inv, pc)
    ref = sums / np.bincount(inv)[:, None]
    assert len(th.cp32) == len(uniq)
    # key order (ix-major, iy-minor) is the order of `uniq`: same rows in the same order
    assert np.allclose(th.cp64, ref, atol=1e-12)
    assert np.allclose(th.cp32, ref, atol=1e-5)


def test_thin_matches_dict_brute_force():
    pos_m, f_m = random_cloud(30, 9, -2.0, 2.0)
    pc, _, _ = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0, alpha=2.0))
    bf = _brute.thin_dict(pc, 0.5)
    assert len(bf) == len(th.cp32)
    for r, (key, members, cen) in enumerate(bf):
        assert tuple(th.key[r, :2]) == key
        assert list(th.order[th.start[r]: th.start[r] + th.length[r]]) == members  # index tie-break
        assert np.allclose(th.cp64[r], cen, atol=1e-12)


def test_thin_fp32_centroid_is_sequential():
    # three members summed left to right in fp32, then divided by float32(3)
    pc = np.array([[0.1, 0.1], [0.2, 0.15], [0.24, 0.05]], np.float32)
    th = v2.thin(pc, v2.Params(tau=1.0))
    s = np.float32(np.float32(np.float32(0.0) + pc[0, 0]) + pc[1, 0])
    s = np.float32(s + pc[2, 0])
    assert len(th.cp32) == 1 and th.cp32[0, 0] == np.float32(s / np.float32(3.0))


def test_refine_ring_known_answers():
    # tests/test_warp_propagation.py:79-112
    posA, dirA, valA = _ring((0.0, 0.0), 0.2, range(7), [0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7])
    posB, dirB, valB = _ring((10.0, 0.0), 0.2, [0, 1, 2], [0.1, 0.2, 0.3])
    pos4, dir4, val4 = _ring((20.0, 0.0), 0.2, [0, 1, 2, 3], [0.1, 0.3, 0.5, 0.7])
    pos5, dir5, val5 = _ring((30.0, 0.0), 0.2, [0, 1, 2, 3, 4], [0.1, 0.3, 0.5, 0.7, 0.9])
    pos = np.vstack([posA, posB, pos4, pos5])
    dirs = np.concatenate([dirA, dirB, dir4, dir5])
    vals = np.concatenate([valA, valB, val4, val5])
    cp = np.array([[0.0, 0.0], [10.0, 0.0], [20.0, 0.0], [30.0, 0.0]])
    new, kappa = v2.refine(cp, pos, vals, dirs, 1.0, 4)
    w = np.exp(valA)
    assert np.allclose(new[0], (w[:, None] * posA).sum(0) / w.sum(), atol=1e-6)
    assert np.allclose(new[1], [10.0, 0.0]) and np.allclose(new[2], [20.0, 0.0])
    assert not np.allclose(new[3], [30.0, 0.0], atol=1e-3)
    assert list(kappa) == [7, 3, 4, 5]


def test_refine_matches_brute_force():
    pos_m, f_m = random_cloud(25, 4, -2.0, 2.0)
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), (0.3 * f_m).astype(np.float32), 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    new, _ = v2.refine(th.cp32, pc, vc, dc, 1.0, 4)
    bf = _brute.refine_loops(th.cp32, pc, vc, dc, 1.0, 4)
    assert np.allclose(new, bf, atol=1e-12)


def test_resample_matches_brute_force_and_balances_mass():
    pos_m, f_m = random_cloud(20, 5, -2.0, 2.0)
    f_m = (0.05 * f_m).astype(np.float32)
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    f_rest = W2
    res = v2.resample(pc, vc, dc, th.cp32, 1.0, f_rest)
    cp_f, out = _brute.resample_loops(pc, vc, dc, th.cp32, 1.0, f_rest)
    assert np.allclose(res.cp_f, cp_f, rtol=1e-12, atol=1e-15)
    assert np.allclose(res.f, out, rtol=1e-12, atol=1e-15)
    # scatter is conservative: delivered mass = component mass - dropped
    assert np.isclose(res.cp_f.sum(), vc.astype(np.float64).sum() - res.dropped, rtol=1e-12)
    assert np.isclose(res.f.sum(), res.cp_f.sum() + res.injected, rtol=1e-12)


def test_gapfill_isolated_point_gets_rest_equilibrium():
    # one component of direction 2 alone: its control point is empty in the other 6 directions;
    # with no other control point in range, sw = dx (self) and f_i = f_rest[i]
    pc = np.array([[3.0, 4.0]], np.float32)
    res = v2.resample(pc, np.array([0.2]), np.array([2]), pc, 1.0, 0.7 * W2)
    assert np.isclose(res.f[0, 2], 0.2)
    others = [0, 1, 3, 4, 5, 6]
    assert np.allclose(res.f[0, others], 0.7 * W2[others])


def test_component_out_of_range_is_dropped():
    pc = np.array([[0.0, 0.0], [5.0, 0.0]], np.float32)
    res = v2.resample(pc, np.array([0.3, 0.4]), np.array([1, 1]), pc[:1], 1.0, W2)
    assert np.isclose(res.dropped, 0.4)
    assert np.isclose(res.cp_f[0, 1], 0.3)


def test_rest_field_preserved_like_reference():
    # tests/test_warp_simulator.py:24-34 on the oracle
    pos, f = workloads.rest_pulse(6)
    sim = v2.Simulator(pos, f, tau=0.6, rho_rest=1.0)
    for _ in range(4):
        sim.step()
    xp, fp = sim.moments()
    rho, _, _ = bin_fields(xp, fp, -3, 3, h=1.0)
    assert abs(rho[rho > 0].mean() - 1.0) < 0.2


def test_periodic_saturated_field_conserves_mass_exactly():
    case = workloads.taylor_green(n=16)
    kw = case.kw
    sim = v2.Simulator(case.pos, case.f, kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"],
                       box=kw["box"])
    for _ in range(3):
        m_in = sim.f_m.astype(np.float64).sum()  # the fp32 state entering the step
        sim.step()
        assert sim.last.res.dropped == 0.0 and sim.last.res.injected == 0.0
        assert len(sim.pos_m) == 16 * 16
        assert abs(sim.last.res.f.sum() / m_in - 1) < 1e-12


def test_periodic_wrap_keeps_points_in_their_cells():
    L, cs = 1.0, 0.125
    x = np.array([-1e-9, 1.0 - 1e-9, 1.0, 0.5, -0.25], np.float32)
    w = v2.wrap32(x, L, cs, 8)
    assert ((w >= 0) & (w < np.float32(L))).all()
    assert (np.floor(w / np.float32(cs)) < 8).all()


def test_two_level_keys():
    prm = v2.Params(tau=1.0, dx=1.0, alpha=4.0, two_level=(0.0, 0.0, 1.0))
    pos = np.array([[0.1, 0.1], [0.3, 0.1], [2.1, 2.1], [2.4, 2.3]], np.float32)
    qx, qy, sub = v2.thin_keys(pos, prm)
    assert list(sub[:2]) == [0, 2]  # inner: fine sub-cells of coarse cell (0, 0)
    assert list(sub[2:]) == [4, 4]  # outer: one coarse cell (4, 4) holds both
    assert list(qx[2:]) == [4, 4] and list(qy[2:]) == [4, 4]
    th = v2.thin(pos, prm)
    assert len(th.cp32) == 3


def test_collide_then_equilibrium_is_fixed_point():
    rng = np.random.default_rng(3)
    rho = rng.uniform(0.5, 2, 10)
    u = rng.uniform(-0.1, 0.1, (10, 2))
    feq = equilibrium(rho, u)
    r = v2.step(np.zeros((1, 2)), feq[:1], v2.Params(tau=0.7))
    assert np.allclose(r.f_star, feq[:1], atol=1e-15)


@pytest.mark.parametrize("n", [8, 12])
def test_step_point_count_periodic(n):
    case = workloads.taylor_green(n=n, r_res=None)
    kw = case.kw
    r = v2.step(case.pos, case.f, v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], box=kw["box"],
                                             rho_rest=kw["rho_rest"]))
    assert len(r.pos) == n * n


# ------------------------------------------------------------ options: periodic box, r_res, two-level
# Pinned by the loop restatements of tests/_brute.py, which find the periodic image by trying all
# three per coordinate and decide inner / outer cells per point; the clouds straddle the seams.

def _seam_cloud(m, seed, L, scale=0.3):
    """m moments of a periodic box [0, L)^2 placed within about dx of the x = 0 and y = 0 seams
    (and their corner), so that many pairs meet only through a periodic image."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.9, 0.9, (m, 2)) % L
    f = rng.uniform(0.01, 1.0, (m, 7)) * scale
    return pos.astype(np.float32), f.astype(np.float32)


def _image_pairs(a, b, r, L):
    """Pairs within r only through a periodic image (the plain difference is farther than r)."""
    n = 0
    for ax, ay in a:
        for bx, by in b:
            if _brute.d32p(ax, ay, bx, by, L) <= np.float32(r) < _brute.d32p(ax, ay, bx, by, None):
                n += 1
    return n


@pytest.mark.parametrize("two_level", [None, (1.5, 1.5, 0.8)])
def test_thin_periodic_matches_dict_brute_force(two_level):
    L = 3.0
    prm = v2.Params(tau=1.0, dx=1.0, box=L, two_level=two_level)
    pos_m, f_m = _seam_cloud(30, 11, L)
    pc, _, _ = v2.disperse(pos_m, f_m, prm.dx, prm)
    th = v2.thin(pc, prm)
    bf = _brute.thin_dict_options(pc, prm.cs, n_period=prm.ncell, two_level=two_level)
    assert len(bf) == len(th.cp32)
    for r, (key, members, cen) in enumerate(bf):
        assert tuple(th.key[r]) == key
        assert list(th.order[th.start[r]: th.start[r] + th.length[r]]) == members
        assert np.allclose(th.cp64[r], cen, atol=1e-12)
    if two_level is not None:
        subs = {k[2] for k, _, _ in bf}
        assert 4 in subs and subs & {0, 1, 2, 3}  # both inner and outer cells present


def test_thin_two_level_open_matches_dict_brute_force():
    prm = v2.Params(tau=1.0, dx=1.0, two_level=(0.2, -0.1, 1.1))
    pos_m, f_m = random_cloud(30, 12, -2.0, 2.0)
    pc, _, _ = v2.disperse(pos_m.astype(np.float32), f_m.astype(np.float32), 1.0, prm)
    th = v2.thin(pc, prm)
    bf = _brute.thin_dict_options(pc, prm.cs, two_level=prm.two_level)
    assert len(bf) == len(th.cp32)
    for r, (key, members, cen) in enumerate(bf):
        assert tuple(th.key[r]) == key
        assert list(th.order[th.start[r]: th.start[r] + th.length[r]]) == members
        assert np.allclose(th.cp64[r], cen, atol=1e-12)
    subs = {k[2] for k, _, _ in bf}
    assert 4 in subs and subs & {0, 1, 2, 3}


def test_refine_periodic_matches_brute_force():
    L = 6.0  # a sparse cloud at the seams: both hard (kappa <= 4) and moved points
    prm = v2.Params(tau=1.0, dx=1.0, box=L)
    pos_m, f_m = _seam_cloud(10, 13, L)
    pc, vc, dc = v2.disperse(pos_m, f_m, prm.dx, prm)
    th = v2.thin(pc, prm)
    new, kappa = v2.refine(th.cp32, pc, vc, dc, prm.dx, 4, prm)
    bf = _brute.refine_loops_periodic(th.cp32, pc, vc, dc, prm.dx, 4, L)
    assert np.abs(new - bf).max() <= 1e-12
    assert ((new >= 0) & (new < float(np.float32(L)))).all()
    assert (kappa > 4).any() and (kappa <= 4).any()
    assert _image_pairs(th.cp32, pc, prm.dx, L) > 0


@pytest.mark.parametrize("periodic,r_res", [(True, None), (True, 0.375), (False, 0.375), (False, 1.5)])
def test_resample_options_match_brute_force(periodic, r_res):
    L = 3.0
    prm = v2.Params(tau=1.0, dx=1.0, box=L if periodic else None, r_res=r_res)
    if periodic:
        pos_m, f_m = _seam_cloud(20, 14, L, scale=0.05)
    else:
        pos_m, f_m = random_cloud(20, 15, -2.0, 2.0)
        pos_m, f_m = pos_m.astype(np.float32), (0.05 * f_m).astype(np.float32)
    pc, vc, dc = v2.disperse(pos_m, f_m, prm.dx, prm)
    th = v2.thin(pc, prm)
    f_rest = 0.7 * W2
    res = v2.resample(pc, vc, dc, th.cp32, prm.dx, f_rest, prm)
    r = 1.0 if r_res is None else r_res  # the resample radius, dx unless given
    cp_f, out = _brute.resample_loops_options(pc, vc, dc, th.cp32, r, f_rest, L=prm.box)
    assert np.allclose(res.cp_f, cp_f, rtol=1e-12, atol=1e-15)
    assert np.allclose(res.f, out, rtol=1e-12, atol=1e-15)
    assert (res.cp_f < prm.eps).any()  # some directions are gap-filled
    assert np.isclose(res.cp_f.sum(), vc.astype(np.float64).sum() - res.dropped, rtol=1e-12)
    if periodic:
        assert _image_pairs(pc, th.cp32, r, L) > 0


# ------------------------------------------------------------------------- soft_outer spawn
def _spawn_ring(dirs, center=(5.0, -3.0)):
    pos, d, _ = _ring(center, 0.2, dirs, [0.1] * len(dirs))
    return pos.astype(np.float32), d


@pytest.mark.parametrize("dirs,normal", [
    ([1, 2, 3, 4, 5], (-0.5, np.sqrt(3) / 2)),  # sum (-1/2, sqrt3/2): unit already
    ([1, 1, 1, 2, 3, 4, 5], (np.sqrt(3) / 2, 0.5)),  # multiplicity counts: (3/2, sqrt3/2) / sqrt3
    ([2, 3, 4, 5, 6], (-1.0, 0.0)),  # c_2 + c_3 + c_4 + c_5 + c_6 = (-1, 0): the c_y cancel
])
def test_spawn_ring_known_normals(dirs, normal):
    # direbm/reference/simulator.py:153-166: p + normal * 2 (1 - sqrt3/2) dx, kappa_hard < kappa <= kappa_soft
    pc, dc = _spawn_ring(dirs)
    cp = np.array([[5.0, -3.0]], np.float32)
    _, kappa = v2.refine(cp, pc, np.zeros(len(pc)), dc, 1.0, 4)
    assert kappa[0] == 5
    par, born = v2.spawn(cp, pc, dc, kappa, 1.0, 4, 5)
    assert list(par) == [0]
    exp = np.array([5.0, -3.0]) + np.array(normal) * (2 * (1 - np.sqrt(3) / 2))
    assert np.allclose(born[0], exp, atol=1e-6)


@pytest.mark.parametrize("dirs,kappa_soft", [
    ([0, 1, 4, 2, 5], 5),  # balanced: c_1 + c_4 = c_2 + c_5 = 0 -> no front, no spawn
    ([1, 2, 3, 4], 5),  # kappa 4: hard outer
    ([1, 2, 3, 4, 5, 6], 5),  # kappa 6: inner
    ([1, 2, 3, 4, 5], 0),  # spawn off
])
def test_spawn_ring_no_spawn(dirs, kappa_soft):
    pc, dc = _spawn_ring(dirs)
    cp = np.array([[5.0, -3.0]], np.float32)
    _, kappa = v2.refine(cp, pc, np.zeros(len(pc)), dc, 1.0, 4)
    par, born = v2.spawn(cp, pc, dc, kappa, 1.0, 4, kappa_soft)
    assert len(par) == 0 and len(born) == 0


def test_spawn_matches_loop_restatement():
    pos_m, f_m = random_cloud(40, 21, -4.0, 4.0)
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), f_m.astype(np.float32), 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    _, kappa = v2.refine(th.cp32, pc, vc, dc, 1.0, 4)
    par, born = v2.spawn(th.cp32, pc, dc, kappa, 1.0, 4, 5)
    bf = _brute.spawn_loops(th.cp32, pc, dc, 1.0, 4, 5, C2, v2.SOFT_OFFSET)
    assert len(bf) == len(par) > 0
    assert list(par) == [r for r, _ in bf]
    assert np.allclose(born, np.array([x for _, x in bf]), atol=1e-6)


def test_step_with_spawn_appends_spawned_points():
    pos_m, f_m = random_cloud(40, 22, -4.0, 4.0)
    prm_off = v2.Params(tau=0.7)
    prm_on = v2.Params(tau=0.7, kappa_soft=5)
    a = v2.step(pos_m.astype(np.float32), (0.2 * f_m).astype(np.float32), prm_off)
    b = v2.step(pos_m.astype(np.float32), (0.2 * f_m).astype(np.float32), prm_on)
    P = len(a.pos)
    assert len(b.pos) > P
    assert np.array_equal(b.pos[:P], a.pos)  # the regular points first, unchanged
    _, born = v2.spawn(b.thin.cp32, *v2.disperse(pos_m.astype(np.float32), b.f_star.astype(np.float32), 1.0)[::2],
                       b.kappa, 1.0, 4, 5)
    assert np.array_equal(b.pos[P:].astype(np.float32), born)
    # the spawned points take part in the resample: mass balance holds over all points
    assert np.isclose(b.f.sum(), b.res.cp_f.sum() + b.res.injected, rtol=1e-12)


def _front_anisotropy(kappa_soft, steps=10, sectors=36):
    # tests/test_reference.py:70-79 on this step: a point pulse's front radius per angular sector
    pos = np.zeros((1, 2), np.float32)
    f = equilibrium(np.float64(1.6), np.zeros(2))[None].astype(np.float32)
    sim = v2.Simulator(pos, f, 0.6, kappa_soft=kappa_soft)
    for _ in range(steps):
        sim.step()
    x, _ = sim.moments()
    r = np.linalg.norm(x, axis=1)
    th = np.arctan2(x[:, 1], x[:, 0]) % (2 * np.pi)
    edges = np.linspace(0, 2 * np.pi, sectors + 1)
    front = np.array([r[(th >= edges[k]) & (th < edges[k + 1])].max(initial=0.0) for k in range(sectors)])
    return front.std() / front.mean()


def test_spawn_reduces_circular_anisotropy_like_reference():
    # tests/test_reference.py:82-85: the spawn exists to counter the hexagonal anisotropy
    assert _front_anisotropy(5) < 0.5 * _front_anisotropy(0)
