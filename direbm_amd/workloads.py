"""Synthetic initial states of the benchmark configurations (host-side numpy generation).

A point carries the mass of the cell it samples: f = f_eq(rho * A, u) with A its cell area
(the dilution convention of direbm/fields.py:3-7), so the binned density field is rho.  The
thinning cell h is the seeding cell; dx = alpha * h with alpha = 4 (direbm/constants.py:46-48).
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ._host import equilibrium


@dataclass
class Case:
    name: str
    pos: np.ndarray  # (M, 2) float32
    f: np.ndarray  # (M, 7) float32
    kw: dict = field(default_factory=dict)  # GpuSimulator keyword arguments (tau, dx, alpha, ...)


def taylor_green(n=128, u0=0.05, tau=0.8, rho=1.0, seed=0, L=2 * np.pi, alpha=4.0, r_res=None) -> Case:
    """Periodic Taylor-Green vortex on [0, L)^2 seeded with n x n jittered cell centres.

    u_x = -u0 cos x sin y, u_y = u0 sin x cos y (pattern of experiments/exp_taylor_green.py:40-42),
    jitter U(-0.1 h, 0.1 h).
    """
    h = L / n
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))
    k = 2 * np.pi / L
    u = np.stack([-u0 * np.cos(k * p[:, 0]) * np.sin(k * p[:, 1]), u0 * np.sin(k * p[:, 0]) * np.cos(k * p[:, 1])], 1)
    f = equilibrium(np.full(n * n, rho * h * h), u)
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=rho * h * h, box=L)
    if r_res is not None:
        kw["r_res"] = r_res
    return Case(f"taylor_green_{n}", p.astype(np.float32), f.astype(np.float32), kw)


def gaussian_pulse_two_level(n=1472, h=0.25, amp=1e-3, sigma_cells=8.0, tau=0.6, r_in_frac=0.3, seed=0,
                             alpha=4.0, two_level=True) -> Case:
    """Gaussian density pulse on rho = 1 at rest, unbounded domain, two-level thinning.

    Square [0, n h)^2; one jittered point per fine cell (h) inside radius r_in_frac * n h of the
    centre and one per coarse cell (2h) outside (jitter U(-0.1, 0.1) of the cell size).
    rho = 1 + amp exp(-r^2 / (2 (sigma_cells h)^2)).  n = 1472 gives ~1.0M points.
    """
    if n % 2:
        raise ValueError("n must be even")
    rng = np.random.default_rng(seed)
    S = n * h
    c = 0.5 * S
    rin = r_in_frac * S
    gf = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(gf, gf, indexing="ij")
    fine = np.stack([X.ravel(), Y.ravel()], 1)
    gc = (np.arange(n // 2) + 0.5) * 2 * h
    X2, Y2 = np.meshgrid(gc, gc, indexing="ij")
    coarse = np.stack([X2.ravel(), Y2.ravel()], 1)
    if two_level:
        # a coarse cell is inner when its centre lies inside the disk: its 4 fine cells are seeded
        outer = ((coarse - c) ** 2).sum(1) >= rin * rin
        inner = ~outer.reshape(n // 2, n // 2).repeat(2, 0).repeat(2, 1).ravel()
        pf = fine[inner] + rng.uniform(-0.1 * h, 0.1 * h, (inner.sum(), 2))
        pc = coarse[outer] + rng.uniform(-0.2 * h, 0.2 * h, (outer.sum(), 2))
        pos = np.vstack([pf, pc])
        area = np.concatenate([np.full(len(pf), h * h), np.full(len(pc), 4 * h * h)])
    else:
        pos = fine + rng.uniform(-0.1 * h, 0.1 * h, fine.shape)
        area = np.full(len(pos), h * h)
    r2 = ((pos - c) ** 2).sum(1)
    rho = 1.0 + amp * np.exp(-r2 / (2 * (sigma_cells * h) ** 2))
    f = equilibrium(rho * area, np.zeros((len(pos), 2)))
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=h * h)
    if two_level:
        kw["two_level"] = (c, c, rin)
    return Case(f"gaussian_pulse_{len(pos)}", pos.astype(np.float32), f.astype(np.float32), kw)


def rest_pulse(half, pulse_rho=None, rho=1.0):
    """Rest field on the integer grid [-half, half]^2 with an optional centre pulse.

    experiments/exp_gpu_bench.py:44-53 and tests/test_warp_simulator.py:11-21.
    """
    f_rest = equilibrium(np.float64(rho), np.zeros(2))
    pos, fs = [], []
    for i in range(-half, half + 1):
        for j in range(-half, half + 1):
            pos.append([float(i), float(j)])
            if pulse_rho is not None and i == 0 and j == 0:
                fs.append(equilibrium(np.float64(pulse_rho), np.zeros(2)))
            else:
                fs.append(f_rest.copy())
    return np.array(pos), np.stack(fs)


def lbm_gaussian_density(ni, nj, amp=1e-3, sigma=8.0):
    """rho = 1 + amp exp(-r^2 / (2 sigma^2)) on the hex lattice, r from the centre node."""
    ii, jj = np.meshgrid(np.arange(ni), np.arange(nj))
    ci, cj = ni // 2, nj // 2
    x = (ii + jj * 0.5) - (ci + cj * 0.5)
    y = (jj - cj) * np.sqrt(3.0) / 2.0
    return 1.0 + amp * np.exp(-(x * x + y * y) / (2 * sigma * sigma))


def shear_layer(n=128, u0=0.05, tau=0.7, rho=1.0, seed=0, L=2 * np.pi, alpha=4.0, delta=0.05, pert=1e-4) -> Case:
    """Periodic double shear layer on [0, L)^2: u_x = u0 tanh((y - L/4)/(delta L)) for y < L/2 and
    u0 tanh((3L/4 - y)/(delta L)) above, plus a seeded U(-pert, pert) perturbation of u_y (the
    periodic counterpart of the shear layer of experiments/exp_shear_3d.py)."""
    h = L / n
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))
    y = p[:, 1]
    ux = np.where(y < 0.5 * L, u0 * np.tanh((y - 0.25 * L) / (delta * L)), u0 * np.tanh((0.75 * L - y) / (delta * L)))
    uy = rng.uniform(-pert, pert, n * n)
    f = equilibrium(np.full(n * n, rho * h * h), np.stack([ux, uy], 1))
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=rho * h * h, box=L)
    return Case(f"shear_layer_{n}", p.astype(np.float32), f.astype(np.float32), kw)


def fourier_velocity(n=512, modes=8, amp=0.03, seed=0, L=2 * np.pi):
    """Smooth periodic velocity field: a sum of `modes` random Fourier modes on [0, L)^2.

    Mode j: integer wavevector k_j with 1 <= |k_j|_inf <= 3, random phase and direction; the
    field is scaled so that its largest speed on the n x n seeding grid is `amp`.  Returns a
    function u(points (M, 2)) -> (M, 2).
    """
    rng = np.random.default_rng(seed)
    ks = []
    while len(ks) < modes:
        k = rng.integers(-3, 4, size=2)
        if np.abs(k).max() >= 1:
            ks.append(k)
    ks = np.array(ks, np.float64) * (2 * np.pi / L)
    phase = rng.uniform(0, 2 * np.pi, modes)
    ang = rng.uniform(0, 2 * np.pi, modes)
    dirs = np.stack([np.cos(ang), np.sin(ang)], 1)

    def raw(p):
        s = np.sin(p @ ks.T + phase)  # (M, modes)
        return s @ dirs

    g = (np.arange(n) + 0.5) * (L / n)
    X, Y = np.meshgrid(g, g, indexing="ij")
    scale = amp / np.abs(raw(np.stack([X.ravel(), Y.ravel()], 1))).max()
    return lambda p: scale * raw(np.asarray(p, np.float64))


def adjoint_case(n=512, seed=0, tau=0.8, amp=0.03, L=2 * np.pi, alpha=4.0, jitter_seed=0):
    """Differentiable-rollout workload: n x n jittered points on the periodic [0, L)^2 at rho = 1,
    initial velocity a sum of 8 random Fourier modes (amplitude amp, velocity seed `seed`).

    Returns (pos (M, 2) float32, rho0 (M,) per-point mass rho h^2, u0 (M, 2), GpuSimulator kw).
    The positions depend on jitter_seed only, so cases with different velocity seeds share them.
    """
    h = L / n
    rng = np.random.default_rng(jitter_seed)
    g = (np.arange(n) + 0.5) * h
    X, Y = np.meshgrid(g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n * n, 2))
    u = fourier_velocity(n, 8, amp, seed, L)(p)
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=h * h, box=L)
    return p.astype(np.float32), np.full(n * n, h * h, np.float32), u.astype(np.float32), kw


# The gap-fill's emptiness threshold, 1e-6 (direbm/warp/propagation.py:243), is absolute: it is
# set for points of unit mass.  A point here carries the mass of its seeding cell, rho h^D, so the
# 3D workloads, whose h^3 brings every population below 1e-6 at a few dozen cells a side, scale
# the threshold with it (eps = 1e-6 h^3).  Left absolute, every direction would read as empty and
# be refilled with rest mass.
_EPS_UNIT = 1e-6


def _equilibrium3(rho, u):
    """D3Q13 equilibrium (FP64) with the fp32 velocity set of the 3D kernels' reference values."""
    phi = (1.0 + np.sqrt(5.0)) / 2.0
    v = np.array([[0, 1, phi], [0, 1, -phi], [0, -1, phi], [0, -1, -phi], [1, phi, 0], [1, -phi, 0],
                  [-1, phi, 0], [-1, -phi, 0], [phi, 0, 1], [phi, 0, -1], [-phi, 0, 1], [-phi, 0, -1]])
    c = np.vstack([np.zeros(3), v / np.linalg.norm(v[0])])
    w = np.array([0.4] + [0.05] * 12)
    rho = np.asarray(rho, np.float64)
    u = np.asarray(u, np.float64)
    cu = u @ c.T
    usq = (u * u).sum(-1)
    return rho[..., None] * w * (1 + 5 * cu + 12.5 * cu * cu - 2.5 * usq[..., None])


def shear_layer_3d(n=32, u0=0.05, tau=0.7, delta=0.05, pert=1e-4, seed=0, L=1.0, alpha=4.0, sphere=None) -> Case:
    """3D periodic double shear layer on [0, L)^3 seeded with n^3 jittered cell centres:
    u_x = u0 tanh((y - L/4)/(delta L)) below L/2 and u0 tanh((3L/4 - y)/(delta L)) above, plus a
    seeded U(-pert, pert) perturbation of u_y (the periodic counterpart of the shear layer of
    experiments/exp_shear_3d.py), rho = 1 (per-point mass h^3), dx = alpha h.

    sphere: radius as a fraction of L of an obstacle at the box centre (fluid outside); seeds
    inside it are removed and kw["sphere"] = (cx, cy, cz, R)."""
    h = L / n
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n ** 3, 3))
    y = p[:, 1]
    ux = np.where(y < 0.5 * L, u0 * np.tanh((y - 0.25 * L) / (delta * L)), u0 * np.tanh((0.75 * L - y) / (delta * L)))
    u = np.stack([ux, rng.uniform(-pert, pert, n ** 3), np.zeros(n ** 3)], 1)
    f = _equilibrium3(np.full(n ** 3, h ** 3), u)
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=h ** 3, box=L, eps=_EPS_UNIT * h ** 3)
    name = f"shear_layer_3d_{n}"
    if sphere is not None:
        c, R = 0.5 * L, float(sphere) * L
        out = ((p - c) ** 2).sum(1) >= R * R
        p, f = p[out], f[out]
        kw["sphere"] = (c, c, c, R)
        name += f"_sphere{sphere:g}"
    return Case(name, p.astype(np.float32), f.astype(np.float32), kw)


def pulse_3d(n=12, h=0.25, amp=0.05, sigma_cells=2.0, tau=0.6, seed=0, alpha=4.0) -> Case:
    """Open-domain 3D Gaussian density pulse on rho = 1 at rest: n^3 jittered points of spacing h."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + rng.uniform(-0.1 * h, 0.1 * h, (n ** 3, 3))
    c = 0.5 * n * h
    r2 = ((p - c) ** 2).sum(1)
    rho = 1.0 + amp * np.exp(-r2 / (2 * (sigma_cells * h) ** 2))
    f = _equilibrium3(rho * h ** 3, np.zeros((n ** 3, 3)))
    kw = dict(tau=tau, dx=alpha * h, alpha=alpha, rho_rest=h ** 3, eps=_EPS_UNIT * h ** 3)
    return Case(f"pulse_3d_{n}", p.astype(np.float32), f.astype(np.float32), kw)


def batch_cases(ninst=256, nmin=4096, nmax=262144, seed=1):
    """Independent periodic instances: point counts log-uniform in [nmin, nmax] (rounded to a
    square n x n), Taylor-Green or shear layer 50/50, tau ~ U(0.55, 1.5), u0 ~ U(0.01, 0.08)."""
    rng = np.random.default_rng(seed)
    cases = []
    for b in range(ninst):
        m = float(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        n = max(8, int(round(np.sqrt(m))))
        tau = float(rng.uniform(0.55, 1.5))
        u0 = float(rng.uniform(0.01, 0.08))
        if rng.uniform() < 0.5:
            cases.append(taylor_green(n=n, u0=u0, tau=tau, seed=seed * 1000 + b))
        else:
            cases.append(shear_layer(n=n, u0=u0, tau=tau, seed=seed * 1000 + b))
    return cases
