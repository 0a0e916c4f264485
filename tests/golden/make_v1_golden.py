"""Write golden vectors of the reference's v1 solver (direbm.reference.Simulator).

The v1 algorithm is order-dependent (greedy control points in dict-of-cells order), so no closed
form pins it; these vectors, written by the reference package itself, pin oracle/v1.py.
Run in a container that holds the reference at /root/reference:

    python tests/golden/make_v1_golden.py

Output: tests/golden/v1_golden.npz (inputs and the moments after each step, in list order).
"""

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, "/root/reference")
from direbm import equilibrium  # noqa: E402
from direbm.reference import Moment, Simulator  # noqa: E402


def cases():
    out = {}
    f_rest = equilibrium(np.float64(1.0), np.zeros(2))
    pos, f = [], []
    for i in range(-3, 4):
        for j in range(-3, 4):
            pos.append([float(i), float(j)])
            f.append(equilibrium(np.float64(1.5), np.zeros(2)) if i == 0 and j == 0 else f_rest.copy())
    out["rest_pulse"] = (np.array(pos), np.stack(f), dict(tau=0.6, rho_rest=1.0), 2)
    out["single_pulse"] = (np.zeros((1, 2)), equilibrium(np.float64(1.6), np.zeros(2))[None, :],
                           dict(tau=0.6, rho_rest=1.0), 3)
    rng = np.random.default_rng(4)
    out["random_cloud_soft_off"] = (rng.uniform(-2.0, 2.0, (30, 2)), 0.1 * rng.uniform(0.01, 1.0, (30, 7)),
                                    dict(tau=0.7, rho_rest=0.5, soft_mode="off"), 1)
    return out


def main():
    data = {}
    for name, (pos, f, kw, steps) in cases().items():
        sim = Simulator([Moment(f=f[k].copy(), x=pos[k].copy()) for k in range(len(pos))], **kw)
        data[f"{name}/pos0"] = pos
        data[f"{name}/f0"] = f
        data[f"{name}/tau"] = np.float64(kw["tau"])
        data[f"{name}/rho_rest"] = np.float64(kw["rho_rest"])
        data[f"{name}/soft_off"] = np.int64(kw.get("soft_mode") == "off")
        data[f"{name}/steps"] = np.int64(steps)
        for t in range(steps):
            sim.step()
            data[f"{name}/pos{t + 1}"] = np.array([m.x for m in sim.moments])
            data[f"{name}/f{t + 1}"] = np.stack([m.f for m in sim.moments])
    out = Path(__file__).with_name("v1_golden.npz")
    np.savez_compressed(out, **data)
    print(out, {k: v.shape for k, v in data.items() if k.endswith(("pos1", "pos2", "pos3"))})


if __name__ == "__main__":
    main()
