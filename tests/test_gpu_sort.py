"""drb_radix_sort_pairs (the replacement of warp.utils.radix_sort_pairs, called at
direbm/warp/propagation.py:138) against numpy's stable argsort: keys ascending, ties in input
order, every value carried with its key.  Random 32-bit keys with duplicates at the size of the
benchmark's component stream (7M), 20-bit keys (3 digit passes, as the step's keys), and sizes
around the 4096-key tile."""

import ctypes as C

import numpy as np
import pytest

from direbm_amd import _lib
from direbm_amd._device import stream_ptr, torch

pytestmark = pytest.mark.gpu


def _sort(keys, vals, bits):
    t = torch()
    dev = t.device("cuda:0")
    n = len(keys)
    k = t.zeros(2 * max(n, 1), dtype=t.int32, device=dev)
    v = t.zeros(2 * max(n, 1), dtype=t.int32, device=dev)
    if n:
        k[:n] = t.as_tensor(keys.view(np.int32)).to(dev)
        v[:n] = t.as_tensor(vals.view(np.int32)).to(dev)
    h = C.c_void_p()
    _lib.call("drb_stage_create", C.byref(h))
    try:
        _lib.call("drb_radix_sort_pairs", h, k.data_ptr(), v.data_ptr(), n, bits, stream_ptr(dev))
        t.cuda.synchronize()
    finally:
        _lib.call("drb_stage_destroy", h)
    return k[:n].cpu().numpy().view(np.uint32), v[:n].cpu().numpy().view(np.uint32)


def _check(keys, bits):
    vals = np.arange(len(keys), dtype=np.uint32)
    gk, gv = _sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, keys[order])
    assert np.array_equal(gv, vals[order])


def test_sort_7m_random_32bit_keys_with_duplicates():
    rng = np.random.default_rng(0)
    distinct = rng.integers(0, 2**32, size=2_000_000, dtype=np.uint64).astype(np.uint32)
    keys = distinct[rng.integers(0, len(distinct), size=7_000_000)]  # about 3.5 copies of each key
    assert len(np.unique(keys)) < len(keys)
    _check(keys, 32)


def test_sort_20bit_keys():
    rng = np.random.default_rng(1)
    _check(rng.integers(0, 2**20, size=7_340_032, dtype=np.uint32), 20)


@pytest.mark.parametrize("n", [0, 1, 2, 63, 4095, 4096, 4097, 70_001])
def test_sort_sizes_around_the_tile(n):
    rng = np.random.default_rng(n)
    _check(rng.integers(0, 300, size=n, dtype=np.uint32), 9)


def test_sort_carries_arbitrary_values():
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 2**12, size=100_000, dtype=np.uint32)
    vals = rng.integers(0, 2**32, size=100_000, dtype=np.uint64).astype(np.uint32)
    gk, gv = _sort(keys, vals, 12)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(gk, keys[order]) and np.array_equal(gv, vals[order])


def test_sort_rejects_bad_bits():
    with pytest.raises(_lib.DrbError, match="INVALID_ARG"):
        _sort(np.zeros(4, np.uint32), np.zeros(4, np.uint32), 33)
