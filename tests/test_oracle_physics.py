"""Pins of the oracle physics, LBM and field binning (CPU).

Known answers and invariants of tests/test_physics.py, tests/test_lattices_3d.py,
tests/test_lbm.py and tests/test_fields.py of the reference, plus closed forms.
"""

import numpy as np

from oracle.constants import C2, C3, CS2_2D, CS2_3D, HEX_OFFSETS, W2, W3
from oracle.fields import bin_fields
from oracle.lbm import HexLBM
from oracle.physics import collide, equilibrium, recover


def test_d2q7_constants():
    assert np.allclose(np.linalg.norm(C2[1:], axis=1), 1.0) and np.allclose(C2[0], 0.0)
    assert np.isclose(W2.sum(), 1.0)
    # isotropic second moment sum W c c = cs^2 I
    assert np.allclose(np.einsum("i,ia,ib->ab", W2, C2, C2), CS2_2D * np.eye(2))


def test_d3q13_moment_identities():
    assert np.allclose(np.linalg.norm(C3[1:], axis=1), 1.0)
    assert np.isclose(W3.sum(), 1.0)
    assert np.allclose(np.einsum("i,ia,ib->ab", W3, C3, C3), CS2_3D * np.eye(3))
    m4 = np.einsum("i,ia,ib,ic,id->abcd", W3, C3, C3, C3, C3)
    d = np.eye(3)
    iso = CS2_3D**2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))
    assert np.allclose(m4, iso)


def test_equilibrium_recovers_its_own_macros():
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.5, 2.0, size=16)
    u = rng.uniform(-0.1, 0.1, size=(16, 2))
    rho_r, u_r = recover(equilibrium(rho, u))
    assert np.allclose(rho_r, rho) and np.allclose(u_r, u, atol=1e-12)


def test_equilibrium_at_rest_is_weighted_density():
    rho = np.array([1.0, 3.0])
    assert np.allclose(equilibrium(rho, np.zeros((2, 2))), rho[:, None] * W2[None, :])


def test_equilibrium_closed_form_single_direction():
    # k = 1, c = (1, 0), u = (0.1, 0): rho W1 (1 + 4*0.1 + 8*0.01 - 2*0.01)
    f = equilibrium(np.array([2.0]), np.array([[0.1, 0.0]]))
    assert np.isclose(f[0, 1], 2.0 / 12.0 * (1 + 0.4 + 0.08 - 0.02))


def test_collision_conserves_and_relaxes():
    rng = np.random.default_rng(1)
    f = rng.uniform(0.01, 1.0, size=(32, 7))
    rho0, u0 = recover(f)
    f1 = collide(f, 0.6)
    rho1, u1 = recover(f1)
    assert np.allclose(rho1, rho0) and np.allclose(rho1[:, None] * u1, rho0[:, None] * u0)
    g = np.zeros((1, 7))
    g[0, 0] = 1.0
    feq = equilibrium(*recover(g))
    assert np.abs(collide(g, 0.6) - feq).sum() < np.abs(g - feq).sum()
    # tau = 1 lands exactly on the equilibrium
    assert np.allclose(collide(f, 1.0), equilibrium(rho0, u0))


def test_zero_density_guard():
    rho, u = recover(np.zeros((1, 7)))
    assert rho[0] == 0 and np.all(u == 0)


def test_lbm_rest_steady_and_mass():
    lbm = HexLBM(21, 21, tau=0.6)
    r0, _ = lbm.macroscopic()
    for _ in range(20):
        lbm.step()
    r1, u1 = lbm.macroscopic()
    assert np.allclose(r1, r0) and np.allclose(u1, 0, atol=1e-12)
    lbm = HexLBM(31, 31, tau=0.7)
    lbm.set_pulse(2.0)
    m0 = lbm.f.sum()
    for _ in range(15):
        lbm.step()
    assert np.isclose(lbm.f.sum(), m0)


def test_lbm_streaming_moves_each_population_one_axial_step():
    lbm = HexLBM(9, 7, tau=1e9)  # collision is (numerically) the identity
    lbm.f[:] = 0.0
    for k in range(7):
        lbm.f[3, 4, k] = 1.0 + k
    lbm.stream()
    for k, (di, dj) in enumerate(HEX_OFFSETS):
        assert lbm.f[(3 + dj) % 7, (4 + di) % 9, k] == 1.0 + k


def test_bin_fields_known_answers():
    # tests/test_fields.py:18-31
    f_one = equilibrium(np.float64(1.0), np.array([0.05, -0.02]))
    xs = np.array([[gx, gy] for gx in range(-3, 4) for gy in range(-3, 4)], float)
    fs = np.tile(f_one, (len(xs), 1))
    rho, u, _ = bin_fields(xs, fs, -3, 4, h=1.0)
    pop = rho > 0
    assert np.allclose(rho[pop], 1.0)
    assert np.allclose(u[pop], [0.05, -0.02], atol=1e-12)
