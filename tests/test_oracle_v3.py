"""oracle/v3.py (3D D3Q13 step) pinned by the lattice identities of the reference's 3D tests
(tests/test_lattices_3d.py:8-46), brute-force loop restatements of every stage on tiny inputs,
and the reference's 3D behaviour tests (tests/test_reference_3d.py: a pulse spreads, a rest block
keeps its density), plus exact conservation on a saturated periodic field."""

import numpy as np

from oracle import v3
from oracle.constants import C3, CS2_3D, W3
from tests import _brute3 as b3

F32 = np.float32


def test_lattice_identities():
    assert C3.shape == (13, 3)
    assert np.allclose(np.linalg.norm(C3[1:], axis=1), 1.0)
    assert abs(W3.sum() - 1.0) < 1e-15
    m2 = np.einsum("i,ia,ib->ab", W3, C3, C3)
    assert np.allclose(m2, CS2_3D * np.eye(3))
    m4 = np.einsum("i,ia,ib,ic,id->abcd", W3, C3, C3, C3, C3)
    d = np.eye(3)
    iso = CS2_3D**2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))
    assert np.allclose(m4, iso)


def test_equilibrium_and_collide_invariants():
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.5, 2.0, 50)
    u = rng.uniform(-0.1, 0.1, (50, 3))
    feq = v3.equilibrium(rho, u)
    assert np.allclose(feq.sum(1), rho)
    assert np.allclose(feq @ C3, rho[:, None] * u)
    f = feq + rng.uniform(0, 0.01, feq.shape)
    g = v3.collide(f, 0.7)
    assert np.allclose(g.sum(1), f.sum(1))
    assert np.allclose(g @ C3, f @ C3)
    assert np.allclose(v3.collide(f, 1.0), v3.equilibrium(f.sum(1), (f @ C3) / f.sum(1)[:, None]))


def _cloud(m, seed, lo=-1.5, hi=1.5):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(lo, hi, (m, 3)).astype(F32)
    f = v3.equilibrium(rng.uniform(0.5, 1.5, m) * 0.05, rng.uniform(-0.05, 0.05, (m, 3)))
    return pos, f


def test_disperse_closed_form():
    pos, f = _cloud(5, 1)
    prm = v3.Params(tau=0.6, dx=0.8)
    pc, val, dirs = v3.disperse(pos, f, prm)
    for i in range(5):
        for k in range(13):
            c = 13 * i + k
            exp = (pos[i] + (C3[k].astype(F32) * F32(0.8)).astype(F32)).astype(F32)
            assert np.array_equal(pc[c], exp) and val[c] == f[i, k] and dirs[c] == k


def test_thin_matches_dict_of_cells():
    pos, f = _cloud(30, 2)
    prm = v3.Params(tau=0.6, dx=1.0)
    pc, _, _ = v3.disperse(pos, f, prm)
    th = v3.thin(pc, prm)
    ref = b3.thin_dict(pc, prm.cs)
    assert len(ref) == len(th.start)
    for r, (key, mem, cen) in enumerate(ref):
        assert tuple(th.key[r]) == key
        assert list(th.order[th.start[r]:th.start[r] + th.length[r]]) == mem
        assert tuple(th.cp32[r]) == cen


def test_refine_and_resample_match_loops():
    pos, f = _cloud(12, 3)
    prm = v3.Params(tau=0.6, dx=1.0, rho_rest=0.05)
    pc, val, dirs = v3.disperse(pos, f, prm)
    th = v3.thin(pc, prm)
    new, kap = v3.refine(th.cp32, pc, val, dirs, prm)
    rn, rk = b3.refine_loops(th.cp32, pc, val, dirs, prm.dx, prm.kappa_hard)
    assert np.array_equal(kap, rk)
    assert np.allclose(new, rn, rtol=0, atol=1e-12)
    cp = new.astype(F32)
    res = v3.resample(pc, val, dirs, cp, prm)
    cpf, out = b3.resample_loops(pc, val, dirs, cp, prm.rr, prm.f_rest())
    assert np.allclose(res.cp_f, cpf, rtol=1e-12, atol=1e-15)
    assert np.allclose(res.f, out, rtol=1e-12, atol=1e-15)
    # scatter balance: delivered mass = component mass - dropped
    assert abs(res.cp_f.sum() - (val.sum() - res.dropped)) <= 1e-12 * val.sum()


def test_pulse_spreads_and_rest_block_keeps_density():
    # tests/test_reference_3d.py:9-40 behaviour on the v2-style 3D step
    f = v3.equilibrium(np.float64(1.5), np.zeros(3))[None, :]
    pos = np.zeros((1, 3), F32)
    prm = v3.Params(tau=0.6)
    for _ in range(3):
        r = v3.step(pos, f, prm)
        pos, f = r.pos.astype(F32), r.f
    assert len(pos) > 1 and np.isfinite(f).all() and (f.sum(1) > 0).all()
    assert np.linalg.norm(pos.astype(np.float64), axis=1).max() > 1.0
    g = np.arange(-2, 3, dtype=np.float64)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F32)
    f = np.tile(v3.equilibrium(np.float64(1.0), np.zeros(3)), (len(pos), 1))
    r = v3.step(pos, f, v3.Params(tau=0.6, rho_rest=1.0))
    x = r.pos
    interior = (np.abs(x) < 1.5).all(axis=1)
    assert abs(r.f[interior].sum() / 27.0 - 1.0) < 0.3


def test_periodic_saturated_field_conserves_mass():
    n, L = 10, 5.0  # dx = 4h = 2 < L / 2
    h = L / n
    g = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    rng = np.random.default_rng(4)
    pos = (pos + rng.uniform(-0.1 * h, 0.1 * h, pos.shape)).astype(F32)
    u = 0.03 * np.stack([np.sin(2 * np.pi * pos[:, 1] / L), np.zeros(len(pos)), np.cos(2 * np.pi * pos[:, 0] / L)], 1)
    f = v3.equilibrium(np.full(len(pos), h**3), u)
    prm = v3.Params(tau=0.8, dx=4 * h, rho_rest=h**3, box=L)
    m0 = f.sum()
    for _ in range(2):
        r = v3.step(pos, f, prm)
        assert r.res.dropped == 0.0
        pos, f = r.pos.astype(F32), r.f
        assert abs(f.sum() - m0) <= 1e-12 * m0 + r.res.injected


# ----------------------------------------------------------------------------- spherical obstacle
def test_sphere_hit_and_normal():
    # tests/test_boundary_3d.py:9-16
    sph = (0.0, 0.0, 0.0, 1.0)
    assert v3.inside32(np.zeros((1, 3)), sph)[0]
    assert not v3.inside32(np.array([[2.0, 0.0, 0.0]]), sph)[0]
    hit, n = v3.ray_hit64((2.0, 0.0, 0.0), (0.0, 0.0, 0.0), sph)
    assert hit and np.allclose(n, [1.0, 0.0, 0.0])


def test_ray_hit_is_the_first_crossing_by_bisection():
    rng = np.random.default_rng(3)
    sph = (0.3, -0.2, 0.1, 1.1)
    c = np.array(sph[:3])
    for _ in range(200):
        x0 = c + rng.normal(size=3) * 2.0
        x1 = c + rng.normal(size=3) * 0.6
        if np.linalg.norm(x0 - c) <= 1.1 or np.linalg.norm(x1 - c) >= 1.1:
            continue
        hit, n = v3.ray_hit64(x0, x1, sph)
        assert hit
        lo, hi = 0.0, 1.0  # x0 outside, x1 inside: bisection on |x(t) - c| - R
        for _ in range(80):
            mid = 0.5 * (lo + hi)
            if np.linalg.norm(x0 + mid * (x1 - x0) - c) > 1.1:
                lo = mid
            else:
                hi = mid
        p = x0 + lo * (x1 - x0)
        assert np.allclose(n, (p - c) / 1.1, atol=1e-9)


def test_split_conserves_and_points_forward():
    # tests/test_boundary_3d.py:19-26
    d = np.array([0.3, -0.7, 0.6])
    d /= np.linalg.norm(d)
    parts = v3.split_nd(d)
    assert abs(sum(w for _, w in parts) - 1.0) < 1e-9
    v = sum(w * C3[i] for i, w in parts)
    assert float(v @ d) > 0.0


def test_head_on_bounce_closed_form():
    # c_k reflected by the normal -c_k is -c_k, itself a lattice direction (the icosahedron's
    # vertices come in antipodal pairs) with five neighbours at c.c' = 1/sqrt(5): weights 1 and
    # 5 x 1/5 normalised to 1/2 and 1/10 each.
    for k in range(1, 13):
        d = v3.reflect64(C3[k], -C3[k])
        assert np.allclose(d, -C3[k], atol=1e-15)
        parts = dict(v3.split_nd(d))
        anti = int(np.argmin(C3[1:] @ C3[k])) + 1
        assert abs(parts[anti] - 0.5) < 1e-12
        others = {j: w for j, w in parts.items() if j != anti}
        assert len(others) == 5 and all(abs(w - 0.1) < 1e-12 for w in others.values())


def _sphere_block(pulse_at=(2, 0, 0)):
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
                f.append(pulse if (i, j, k) == pulse_at else f_rest)
    return np.array(pos, F32), np.array(f), sph


def test_sphere_keeps_fluid_out_3d():
    # tests/test_boundary_3d.py:29-46 on the 3D step
    pos, f, sph = _sphere_block()
    prm = v3.Params(tau=0.6, rho_rest=1.0, sphere=sph)
    r = v3.step(pos, f, prm)
    assert len(r.pos) > 0
    assert not v3.inside32(r.pos, sph).any()
    assert np.isfinite(r.f).all()


def test_bounce_conserves_component_mass():
    pos, f, sph = _sphere_block()
    prm = v3.Params(tau=0.6, rho_rest=1.0, sphere=sph)
    pc, val, dirs = v3.disperse(pos, f, prm)
    assert abs(val.sum() - f.sum()) <= 1e-12 * f.sum()
    assert not v3.inside32(pc, sph).any()
    assert len(pc) > 13 * len(pos)  # some components were split
