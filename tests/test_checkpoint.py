"""Checkpoint format (.npz, no pickle): round trip of state, iteration and parameters."""

import numpy as np
import pytest

from direbm_amd import checkpoint


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 1, (10, 2)).astype(np.float32)
    f = rng.uniform(0, 1, (10, 7)).astype(np.float32)
    kws = [dict(tau=0.6, dx=1.0, box=2.0, two_level=(0.5, 0.5, 0.2)), dict(tau=0.9, r_res=1.5)]
    p = tmp_path / "c.npz"
    checkpoint.save(p, pos, f, 17, kws, [4, 6])
    pos2, f2, it, kws2, counts = checkpoint.load(p)
    assert np.array_equal(pos2, pos) and np.array_equal(f2, f) and it == 17
    assert kws2 == kws and list(counts) == [4, 6]
    with np.load(p, allow_pickle=False) as z:  # plain arrays only
        assert set(z.files) == {"format", "pos", "f", "counts", "iteration", "params"}


def test_rejects_inconsistent_counts(tmp_path):
    with pytest.raises(ValueError):
        checkpoint.save(tmp_path / "c.npz", np.zeros((3, 2)), np.zeros((3, 7)), 0, [{}], [2])
