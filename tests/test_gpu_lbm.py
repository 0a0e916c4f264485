"""HIP hex LBM vs the FP64 oracle (restates tests/test_warp_lbm.py:9-39)."""

import numpy as np
import pytest

import direbm_amd as drb
from direbm_amd import workloads
from oracle.lbm import HexLBM
from tests._util import rel_l2

pytestmark = pytest.mark.gpu


def test_gpu_lbm_matches_cpu():
    cpu = HexLBM(21, 21, tau=0.6)
    cpu.set_pulse(1.5)
    gpu = drb.GpuHexLBM(21, 21, tau=0.6)
    gpu.set_pulse(1.5)
    for _ in range(10):
        cpu.step()
        gpu.step()
    rc, uc = cpu.macroscopic()
    rg, ug = gpu.macroscopic()
    assert np.allclose(rc, rg, atol=1e-3, rtol=1e-3)
    assert rel_l2(rg, rc) <= 1e-6
    assert rel_l2(gpu.f.numpy(), cpu.f) <= 1e-6


def test_gpu_lbm_multi_step_call_and_odd_sizes():
    cpu = HexLBM(37, 23, tau=0.7)
    cpu.set_pulse(2.0)
    gpu = drb.GpuHexLBM(37, 23, tau=0.7)
    gpu.set_pulse(2.0)
    for _ in range(7):
        cpu.step()
    gpu.step(7)
    assert rel_l2(gpu.f.numpy(), cpu.f) <= 1e-6


@pytest.mark.parametrize("ni,nj", [(4, 5), (8, 3), (12, 7), (64, 2)])
def test_gpu_lbm_four_node_rows(ni, nj):
    """Rows of a multiple of 4 nodes take the four-nodes-per-thread kernel; at ni = 4 one thread's
    nodes reach both ends of the row, so both wrapped writes (di = +1 at the end, -1 at the start)
    come from the same vector."""
    rng = np.random.default_rng(ni * 100 + nj)
    rho = 1.0 + 0.05 * rng.standard_normal((nj, ni))
    cpu = HexLBM(ni, nj, tau=0.65)
    cpu.set_density(rho)
    gpu = drb.GpuHexLBM(ni, nj, tau=0.65)
    gpu.set_density(rho)
    for _ in range(10):
        cpu.step()
    gpu.step(10)
    assert rel_l2(gpu.f.numpy(), cpu.f) <= 1e-6


def test_gpu_lbm_gaussian_density_100_steps():
    rho = workloads.lbm_gaussian_density(128, 96, amp=1e-3, sigma=8.0)
    cpu = HexLBM(128, 96, tau=0.6)
    cpu.set_density(rho)
    gpu = drb.GpuHexLBM(128, 96, tau=0.6)
    gpu.set_density(rho)
    for _ in range(100):
        cpu.step()
    gpu.step(100)
    rc, uc = cpu.macroscopic()
    rg, ug = gpu.macroscopic()
    # density within the 100-step bound (1e-4); fp32 rounding drifts the mean by ~1e-6, so the
    # 1e-3 perturbation itself is compared after removing the mean
    assert rel_l2(rg, rc) <= 1e-5
    assert rel_l2(rg - rg.mean(), rc - rc.mean()) <= 1e-2
    assert rel_l2(ug, uc) <= 1e-2


def test_gpu_lbm_rest_is_steady():
    gpu = drb.GpuHexLBM(15, 15, tau=0.6, rho0=1.0)
    r0, _ = gpu.macroscopic()
    for _ in range(20):
        gpu.step()
    r1, u1 = gpu.macroscopic()
    assert np.allclose(r1, r0, atol=1e-4)
    assert np.allclose(u1, 0.0, atol=1e-5)


def test_gpu_lbm_conserves_mass():
    gpu = drb.GpuHexLBM(25, 25, tau=0.7)
    gpu.set_pulse(2.0)
    m0 = float(gpu.f.numpy().astype(np.float64).sum())
    for _ in range(15):
        gpu.step()
    assert np.isclose(float(gpu.f.numpy().astype(np.float64).sum()), m0, rtol=1e-6)
