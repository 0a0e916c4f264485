"""Checkpoint format (.npz, no pickle): round trip of state, iteration, parameters and tiles."""

import numpy as np
import pytest

from direbm_amd import checkpoint


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, 1, (10, 2)).astype(np.float32)
    f = rng.uniform(0, 1, (10, 7)).astype(np.float32)
    kws = [dict(tau=0.6, dx=1.0, box=2.0, two_level=(0.5, 0.5, 0.2)), dict(tau=0.9, r_res=1.5)]
    p = tmp_path / "c.npz"
    checkpoint.save(p, pos, f, 17, kws, [4, 6], lanes=(2, 8), depth=(2, 4))
    pos2, f2, it, kws2, counts, lanes, depth = checkpoint.load(p)
    assert np.array_equal(pos2, pos) and np.array_equal(f2, f) and it == 17
    assert kws2 == kws and list(counts) == [4, 6] and lanes == (2, 8) and depth == (2, 4)
    with np.load(p, allow_pickle=False) as z:  # plain arrays only
        assert set(z.files) == {"format", "pos", "f", "counts", "iteration", "params", "lanes", "depth"}


def test_format_1_loads_with_default_lanes(tmp_path):
    """Files written before the lanes were recorded resume with the engine's default lanes."""
    import json

    p = tmp_path / "old.npz"
    with open(p, "wb") as fh:
        np.savez(fh, format=np.int64(1), pos=np.zeros((2, 2), np.float32), f=np.zeros((2, 7), np.float32),
                 counts=np.array([2]), iteration=np.int64(3), params=np.array(json.dumps([{"tau": 0.7}])))
    *_, lanes, depth = checkpoint.load(p)
    assert lanes == checkpoint.DEFAULT_LANES and depth == checkpoint.DEFAULT_DEPTH


def test_format_2_loads_with_default_depth(tmp_path):
    """Files written before the depths were recorded resume with their lanes and the default depths."""
    import json

    p = tmp_path / "old2.npz"
    with open(p, "wb") as fh:
        np.savez(fh, format=np.int64(2), pos=np.zeros((2, 2), np.float32), f=np.zeros((2, 7), np.float32),
                 counts=np.array([2]), iteration=np.int64(3), params=np.array(json.dumps([{"tau": 0.7}])),
                 lanes=np.array([8, 2], np.int32))
    *_, lanes, depth = checkpoint.load(p)
    assert lanes == (8, 2) and depth == checkpoint.DEFAULT_DEPTH


def test_rejects_inconsistent_counts(tmp_path):
    with pytest.raises(ValueError):
        checkpoint.save(tmp_path / "c.npz", np.zeros((3, 2)), np.zeros((3, 7)), 0, [{}], [2])
