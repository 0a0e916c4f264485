"""Write the fixtures of the long-horizon band tests (tests/test_gpu_long_horizon.py).

tg100_band.npz -- config 1 at full size: the periodic Taylor-Green vortex seeded on 128 x 128 jittered cells
(16,384 points, u0 0.05, tau 0.8, resample radius 1.5 h), 100 steps of the FP64 oracle
(oracle/v2.py).  Beside the base run, an ensemble of eight oracle runs whose fp32 state is perturbed
after every step by one ulp -- in one coordinate of 10% of the positions and in 10% of the
distribution values, random signs, a different seed per member -- the size of the rounding
differences between an fp32 engine and the FP64 oracle.  The band of a quantity at step t is the
largest deviation of a member from the base run.

pulse30_band.npz -- the open two-level Gaussian pulse (the benchmark's shape) at 96 x 96 seeding
cells (4,248 points, amplitude 0.05, sigma 4 h), 30 steps: the same base run and ensemble, with the
per-step point count, FP64 mass, and the mass each step dropped (components with no control point
in range) and injected (gap-fill).  The pulse grows at the open boundary every step (the
reference algorithm's own growth), so its ensemble spreads geometrically.

Run in this repository (CPU, a few minutes with one process per run):

    python tests/golden/make_tg100_band.py [tg|pulse]

Output: tests/golden/tg100_band.npz and tests/golden/pulse30_band.npz.
"""

import sys
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from direbm_amd import workloads  # noqa: E402
from oracle import v2  # noqa: E402
from tests._tg import observables  # noqa: E402

N, STEPS, MEMBERS = 128, 100, 8
RECORD = (10, 50, 100)


def case():
    c = workloads.taylor_green(n=N, tau=0.8)
    h = c.kw["dx"] / c.kw["alpha"]
    return workloads.taylor_green(n=N, tau=0.8, r_res=1.5 * h)


def _ulp(a, frac, rng):
    """a with one ulp added or removed in a random `frac` of its elements (fp32)."""
    a = a.copy()
    flat = a.reshape(-1)
    idx = rng.choice(flat.size, size=int(frac * flat.size), replace=False)
    sgn = rng.integers(0, 2, size=idx.size) * 2 - 1
    flat[idx] = np.nextafter(flat[idx], np.where(sgn > 0, np.inf, -np.inf).astype(np.float32))
    return a


def run(member):
    c = case()
    kw = c.kw
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], box=kw["box"],
                    r_res=kw["r_res"])
    rng = np.random.default_rng(100 + member)
    pos, f = c.pos.copy(), c.f.copy()
    series = {"amp": [], "mass": [], "count": []}
    fields = {}
    for t in range(1, STEPS + 1):
        r = v2.step(pos, f, prm)
        pos, f = r.pos.astype(np.float32), r.f.astype(np.float32)
        if member > 0:
            # one ulp in one coordinate of 10% of the points (wrapped back into the box)
            pos = _ulp(pos, 0.05, rng)
            n = prm.ncell
            pos = np.stack([v2.wrap32(pos[:, 0], prm.box, prm.cs, n), v2.wrap32(pos[:, 1], prm.box, prm.cs, n)], 1)
            f = _ulp(f, 0.1, rng)
        ob = observables(pos, f, float(kw["box"]))
        series["amp"].append(ob["amp"])
        series["mass"].append(ob["mass"])
        series["count"].append(len(pos))
        if t in RECORD:
            fields[t] = ob
        print(f"member {member} step {t} amp {ob['amp']:.6e} mass {ob['mass']:.12e} points {len(pos)}", flush=True)
    return member, series, fields


def pulse_case():
    return workloads.gaussian_pulse_two_level(n=96, amp=0.05, sigma_cells=4.0)


PULSE_STEPS = 30


def run_pulse(member):
    c = pulse_case()
    kw = c.kw
    prm = v2.Params(tau=kw["tau"], dx=kw["dx"], alpha=kw["alpha"], rho_rest=kw["rho_rest"], two_level=kw["two_level"])
    rng = np.random.default_rng(200 + member)
    pos, f = c.pos.copy(), c.f.copy()
    rows = []
    for t in range(1, PULSE_STEPS + 1):
        r = v2.step(pos, f, prm)
        pos, f = r.pos.astype(np.float32), r.f.astype(np.float32)
        rows.append([len(pos), float(f.astype(np.float64).sum()), r.res.dropped, r.res.injected])
        if member > 0:
            pos = _ulp(pos, 0.05, rng)
            f = _ulp(f, 0.1, rng)
        print(f"pulse member {member} step {t} {rows[-1]}", flush=True)
    return member, np.array(rows)


def main_pulse():
    c = pulse_case()
    with Pool(MEMBERS + 1) as pool:
        res = sorted(pool.map(run_pulse, range(MEMBERS + 1)), key=lambda r: r[0])
    out = {"members": np.int64(MEMBERS), "pos0_sum": np.float64(c.pos.astype(np.float64).sum()),
           "f0_sum": np.float64(c.f.astype(np.float64).sum())}
    for member, rows in res:
        out["base" if member == 0 else f"m{member}"] = rows  # columns: count, mass, dropped, injected
    np.savez_compressed(ROOT / "tests" / "golden" / "pulse30_band.npz", **out)
    print("wrote tests/golden/pulse30_band.npz")


def main():
    c = case()
    with Pool(MEMBERS + 1) as pool:
        res = sorted(pool.map(run, range(MEMBERS + 1)), key=lambda r: r[0])
    out = {"steps": np.array(RECORD), "n": np.int64(N), "members": np.int64(MEMBERS),
           "pos0_sum": np.float64(c.pos.astype(np.float64).sum()), "f0_sum": np.float64(c.f.astype(np.float64).sum())}
    for member, series, fields in res:
        tag = "base" if member == 0 else f"m{member}"
        out[f"{tag}_amp"] = np.array(series["amp"])
        out[f"{tag}_mass"] = np.array(series["mass"])
        out[f"{tag}_count"] = np.array(series["count"])
        for t in RECORD:
            for q in ("rho", "ux", "uy"):
                out[f"{tag}_{q}_{t}"] = fields[t][q]
    np.savez_compressed(ROOT / "tests" / "golden" / "tg100_band.npz", **out)
    print("wrote tests/golden/tg100_band.npz")


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "both"
    if which in ("tg", "both"):
        main()
    if which in ("pulse", "both"):
        main_pulse()
