"""Host-side logic: workload generation and parameter packing (CPU)."""

import numpy as np

from direbm_amd import workloads
from direbm_amd._host import equilibrium, recover
from direbm_amd.simulator import make_params
from oracle.fields import bin_fields


def test_taylor_green_case():
    c = workloads.taylor_green(n=32)
    assert c.pos.shape == (1024, 2) and c.f.shape == (1024, 7)
    assert (c.pos >= 0).all() and (c.pos < c.kw["box"]).all()
    h = 2 * np.pi / 32
    assert np.isclose(c.kw["dx"], 4 * h)
    rho, u, _ = bin_fields(c.pos, c.f, 0, 2 * np.pi - 1e-9, 4 * h)
    assert np.allclose(rho, 1.0, atol=1e-6)  # one point of mass h^2 per cell of area h^2


def test_gaussian_pulse_two_level_case():
    c = workloads.gaussian_pulse_two_level(n=1472)
    assert 990_000 < len(c.pos) < 1_010_000
    c = workloads.gaussian_pulse_two_level(n=128)
    S = 128 * 0.25
    rho, _, _ = bin_fields(c.pos, c.f, 0, S - 1e-9, 1.0)  # 4h bins hold whole coarse cells
    assert np.allclose(rho, 1.0, atol=2e-3)
    cx, cy, rin = c.kw["two_level"]
    assert np.isclose(cx, S / 2) and np.isclose(rin, 0.3 * S)


def test_make_params():
    p = make_params(0.6, dx=2.0, alpha=4.0, r_res=0.75, box=8.0, two_level=(1.0, 2.0, 3.0))
    assert p.periodic == 1 and p.two_level == 1 and np.isclose(p.r_res, 0.75) and np.isclose(p.r_in, 3.0)
    q = make_params(0.6)
    assert q.periodic == 0 and q.r_res == 0.0 and q.kappa_hard == 4


def test_host_equilibrium_recover_roundtrip():
    rng = np.random.default_rng(0)
    rho = rng.uniform(0.5, 2, 8)
    u = rng.uniform(-0.1, 0.1, (8, 2))
    r, v = recover(equilibrium(rho, u))
    assert np.allclose(r, rho) and np.allclose(v, u)
