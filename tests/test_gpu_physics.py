"""HIP collide / equilibrium vs the FP64 oracle (restates tests/test_warp_physics.py:13-42)."""

import numpy as np
import pytest

import direbm_amd as drb
from oracle.physics import collide as ref_collide
from oracle.physics import equilibrium as ref_equilibrium
from oracle.physics import recover
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def test_equilibrium_matches_oracle():
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.5, 2.0, size=128)
    u = rng.uniform(-0.1, 0.1, size=(128, 2))
    ref = ref_equilibrium(rho.astype(np.float32), u.astype(np.float32))
    got = drb.equilibrium(rho, u)
    assert got.dtype == np.float32 and got.shape == (128, 7)
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5)  # the reference's own tolerance
    assert rel_l2(got, ref) <= 1e-6


def test_collide_matches_oracle():
    rng = np.random.default_rng(1)
    f = rng.uniform(0.01, 1.0, size=(256, 7)).astype(np.float32)
    ref = ref_collide(f, 0.6)
    got = drb.collide(f, 0.6)
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-5)
    assert rel_l2(got, ref) <= 1e-6


def test_collide_conserves_mass_and_momentum():
    rng = np.random.default_rng(2)
    f = rng.uniform(0.01, 1.0, size=(128, 7)).astype(np.float32)
    rho0, u0 = recover(f)
    got = drb.collide(f, 0.6)
    rho1, u1 = recover(got)
    assert np.allclose(rho1, rho0, rtol=1e-6, atol=1e-6)
    assert np.allclose(rho1[:, None] * u1, rho0[:, None] * u0, atol=1e-6)


def test_collide_large_batch_and_zero_density():
    rng = np.random.default_rng(3)
    f = rng.uniform(0.0, 1e-3, size=(100_003, 7)).astype(np.float32)
    f[17] = 0.0  # rho == 0: u = 0, f stays 0 (direbm/physics.py:24 guard)
    got = drb.collide(f, 0.75)
    ref = ref_collide(f, 0.75)
    assert np.isfinite(got).all()
    assert np.all(got[17] == 0.0)
    assert rel_l2(got, ref) <= 1e-6


def test_device_string():
    assert drb.default_device() in ("cpu", "cuda:0")
