"""oracle/v1.py (the timed "direbm/reference" CPU baseline) against golden vectors written by the
reference package itself (tests/golden/make_v1_golden.py), plus the reference's own v1 known
answers (tests/test_reference.py)."""

from pathlib import Path

import numpy as np
import pytest

from oracle import v1
from oracle.fields import bin_fields

GOLD = np.load(Path(__file__).parent / "golden" / "v1_golden.npz")
CASES = sorted({k.split("/")[0] for k in GOLD.files})


@pytest.mark.parametrize("name", CASES)
def test_v1_matches_reference_golden(name):
    g = lambda k: GOLD[f"{name}/{k}"]  # noqa: E731
    sim = v1.Simulator(g("pos0"), g("f0"), float(g("tau")), rho_rest=float(g("rho_rest")),
                       soft_mode="off" if int(g("soft_off")) else "spawn")
    for t in range(int(g("steps"))):
        sim.step()
        x, f = sim.moments()
        assert x.shape == g(f"pos{t + 1}").shape, f"step {t + 1}: moment count"
        assert np.allclose(x, g(f"pos{t + 1}"), rtol=0, atol=1e-12)
        assert np.allclose(f, g(f"f{t + 1}"), rtol=1e-12, atol=1e-15)


def test_v1_rest_field_preserved():
    # tests/test_reference.py: rest stays rest macroscopically
    f_rest = np.array([0.5] + [1 / 12] * 6)
    pos = np.array([[i, j] for i in range(-4, 5) for j in range(-4, 5)], float)
    sim = v1.Simulator(pos, np.tile(f_rest, (len(pos), 1)), 0.6)
    for _ in range(2):
        sim.step()
    x, f = sim.moments()
    rho, _, _ = bin_fields(x, f, -2, 2, 1.0)
    assert abs(rho[rho > 0].mean() - 1.0) < 0.2
