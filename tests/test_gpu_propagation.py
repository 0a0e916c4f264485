"""HIP dispersion / thinning / refine / resample vs the oracle.

Restates tests/test_warp_propagation.py:43-112 on this engine's API and adds bit-exact parity
of cell-thinning (count, order, positions) and FP64-oracle parity of refine and resample.
"""

import numpy as np
import pytest

import direbm_amd as drb
from oracle import v2
from oracle.constants import C2, W2
from tests._util import random_cloud, rel_l2

pytestmark = pytest.mark.gpu


def _ring(center, radius, dirs, vals):
    ang = np.linspace(0.0, 2 * np.pi, len(dirs), endpoint=False)
    pos = np.array(center) + radius * np.c_[np.cos(ang), np.sin(ang)]
    return pos, np.array(dirs, dtype=np.int32), np.array(vals, dtype=np.float64)


def _cell_thin_ref(pos_c, cs):
    # tests/test_warp_propagation.py:27-36
    cells = np.floor(pos_c / cs).astype(np.int64)
    keys = cells[:, 0] * 1_000_003 + cells[:, 1]
    uniq, inv = np.unique(keys, return_inverse=True)
    sums = np.zeros((len(uniq), 2))
    cnt = np.zeros(len(uniq))
    np.add.at(sums, inv, pos_c)
    np.add.at(cnt, inv, 1)
    return sums / cnt[:, None]


def _sorted_rows(a):
    return a[np.lexsort((a[:, 1], a[:, 0]))]


def test_dispersion_matches_formula():
    pos_m, f_m = random_cloud(20, 0)
    pos_c, val_c, dir_c = drb.disperse(pos_m, f_m, dx=1.0)
    expected = (pos_m[:, None, :] + C2[None, :, :] * 1.0).reshape(-1, 2)
    assert np.allclose(pos_c.numpy(), expected, atol=1e-5)
    assert np.allclose(val_c.numpy(), f_m.reshape(-1), atol=1e-6)
    assert np.array_equal(dir_c.numpy(), np.tile(np.arange(7), 20))
    # bit-exact against the fp32 reference formula
    pc_ref, _, _ = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    assert np.array_equal(pos_c.numpy(), pc_ref)


def test_cell_thinning_matches_reference_restatement():
    pos_m, f_m = random_cloud(50, 1)
    pos_c, _, _ = drb.disperse(pos_m, f_m, dx=1.0)
    pc = pos_c.numpy()
    cp, p = drb.create_control_points(pos_c, 0.25)
    ref = _cell_thin_ref(pc, 0.25)
    assert p == len(ref)
    assert np.allclose(_sorted_rows(cp.numpy()), _sorted_rows(ref), atol=1e-4)


@pytest.mark.parametrize("m,seed,cs", [(50, 1, 0.25), (2000, 5, 0.25), (20000, 6, 0.1), (300, 7, 2.0)])
def test_cell_thinning_bit_exact(m, seed, cs):
    pos_m, f_m = random_cloud(m, seed, -7.0, 9.0)
    pos_c, _, _ = drb.disperse(pos_m, f_m, dx=1.0)
    pc = pos_c.numpy()
    cp, p = drb.create_control_points(pos_c, cs)
    th = v2.thin(pc, v2.Params(tau=1.0, dx=1.0, alpha=1.0 / cs))
    assert p == len(th.cp32)
    # same order (ascending reference key) and the same fp32 bits
    assert np.array_equal(cp.numpy(), th.cp32)


def test_thinning_reduces_count_and_is_bounded():
    pos_m, f_m = random_cloud(40, 2, -3.0, 3.0)
    pos_c, _, _ = drb.disperse(pos_m, f_m, dx=1.0)
    _, p = drb.create_control_points(pos_c, cs=0.25)
    assert 0 < p <= pos_c.shape[0]


def test_refine_hard_keeps_inner_moves():
    posA, dirA, valA = _ring((0.0, 0.0), 0.2, [0, 1, 2, 3, 4, 5, 6], [0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7])
    posB, dirB, valB = _ring((10.0, 0.0), 0.2, [0, 1, 2], [0.1, 0.2, 0.3])
    pos = np.vstack([posA, posB])
    dirs = np.concatenate([dirA, dirB])
    vals = np.concatenate([valA, valB])
    cp = np.array([[0.0, 0.0], [10.0, 0.0]])
    new = drb.refine_control_points(cp, pos, vals, dirs, dx=1.0, kappa_hard=4).numpy()
    w = np.exp(valA)
    expected_a = (w[:, None] * posA).sum(0) / w.sum()
    assert np.allclose(new[0], expected_a, atol=1e-4)
    assert np.allclose(new[1], [10.0, 0.0], atol=1e-5)


def test_refine_kappa_threshold():
    pos4, dir4, val4 = _ring((0.0, 0.0), 0.2, [0, 1, 2, 3], [0.1, 0.3, 0.5, 0.7])
    pos5, dir5, val5 = _ring((20.0, 0.0), 0.2, [0, 1, 2, 3, 4], [0.1, 0.3, 0.5, 0.7, 0.9])
    pos = np.vstack([pos4, pos5])
    dirs = np.concatenate([dir4, dir5])
    vals = np.concatenate([val4, val5])
    cp = np.array([[0.0, 0.0], [20.0, 0.0]])
    new = drb.refine_control_points(cp, pos, vals, dirs, dx=1.0, kappa_hard=4).numpy()
    assert np.allclose(new[0], [0.0, 0.0], atol=1e-5)
    assert not np.allclose(new[1], [20.0, 0.0], atol=1e-3)


@pytest.mark.parametrize("m,seed", [(60, 11), (3000, 12)])
def test_refine_matches_oracle(m, seed):
    pos_m, f_m = random_cloud(m, seed, -6.0, 6.0)
    f_m = f_m * 0.2
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), f_m.astype(np.float32), 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    ref, kappa = v2.refine(th.cp32, pc, vc, dc, 1.0, 4)
    got = drb.refine_control_points(th.cp32, pc, vc, dc, dx=1.0, kappa_hard=4).numpy()
    assert np.abs(got - ref).max() <= 2e-5
    assert (kappa > 4).any() and (kappa <= 4).any()


@pytest.mark.parametrize("m,seed", [(60, 21), (3000, 22)])
def test_resample_matches_oracle(m, seed):
    pos_m, f_m = random_cloud(m, seed, -6.0, 6.0)
    f_m = (f_m * 0.05).astype(np.float32)
    pc, vc, dc = v2.disperse(pos_m.astype(np.float32), f_m, 1.0)
    th = v2.thin(pc, v2.Params(tau=1.0))
    cp = th.cp32
    f_rest = (1.0 * W2).astype(np.float32)
    ref = v2.resample(pc, vc, dc, cp, 1.0, f_rest)
    out_pos, out_f = drb.resample(pc, vc, dc, cp, 1.0, f_rest)
    assert np.array_equal(out_pos.numpy(), cp)
    got = out_f.numpy()
    assert rel_l2(got, ref.f) <= 1e-5
    assert np.abs(got - ref.f).max() <= 1e-5 * np.abs(ref.f).max()
    # The same directions were gap-filled.  With a negative rest distribution every gap-filled
    # output, (r / sw) f_rest with sw >= r (the point itself), is negative, and every scattered one
    # is a sum of positive component values, so the sign of an output marks an empty direction.
    empty = ref.cp_f < v2.EPS_EMPTY
    assert empty.any() and (~empty).any()
    _, neg_f = drb.resample(pc, vc, dc, cp, 1.0, -f_rest)
    assert np.array_equal(neg_f.numpy() < 0, empty)
