"""drb_create_control_points_noalloc: the allocation-free thinning entry point, called through the
C-ABI, against drb_create_control_points on the same inputs (bit-exact count and positions), on
two calls in a row, the second with a smaller input; its size limits; and no device memory taken
by a call."""

import ctypes as C

import numpy as np
import pytest

from direbm_amd import _lib
from direbm_amd._device import stream_ptr, torch
from oracle import v2
from tests._util import random_cloud

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20


def _components(m, seed, lo, hi):
    pos_m, f_m = random_cloud(m, seed, lo, hi)
    pc, _, _ = v2.disperse(pos_m.astype(np.float32), f_m.astype(np.float32), 1.0)
    return np.ascontiguousarray(pc, np.float32)


def _thin(fn, pc, cs):
    t = torch()
    dev = t.device("cuda:0")
    d_pc = t.as_tensor(pc).to(dev)
    out = t.full((len(pc), 2), float("nan"), dtype=t.float32, device=dev)
    p = C.c_int64(-1)
    if fn == "stage":
        h = C.c_void_p()
        _lib.call("drb_stage_create", C.byref(h))
        try:
            _lib.call("drb_create_control_points", h, d_pc.data_ptr(), len(pc), C.c_float(cs), out.data_ptr(),
                      len(pc), C.byref(p), stream_ptr(dev))
            t.cuda.synchronize()
        finally:
            _lib.call("drb_stage_destroy", h)
    else:
        _lib.call("drb_create_control_points_noalloc", d_pc.data_ptr(), len(pc), C.c_float(cs), out.data_ptr(),
                  len(pc), C.byref(p), stream_ptr(dev))
        t.cuda.synchronize()
    return out[: p.value].cpu().numpy(), int(p.value)


def test_noalloc_matches_the_stage_entry_point_twice_then_smaller():
    cs = 0.25
    big = _components(30_000, 1, -40.0, 40.0)
    small = _components(700, 2, -3.0, 9.0)
    for pc in (big, small):  # the second call reuses the array the first one left
        ref, pr = _thin("stage", pc, cs)
        got, pg = _thin("noalloc", pc, cs)
        assert pg == pr and pr > 0
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
        # and the oracle's control points (count, order, fp32 bits)
        th = v2.thin(pc, v2.Params(tau=1.0, dx=4 * cs))
        assert len(th.cp32) == pg and np.array_equal(got.view(np.uint32), th.cp32.view(np.uint32))


def test_noalloc_refuses_inputs_beyond_its_array():
    t = torch()
    dev = t.device("cuda:0")
    p = C.c_int64(0)
    # too many components (44 bytes each): refused before any kernel reads them
    n = SCRATCH // 44 + 1
    pc = t.zeros((n, 2), dtype=t.float32, device=dev)
    out = t.empty((n, 2), dtype=t.float32, device=dev)
    with pytest.raises(_lib.DrbError, match="CAPACITY"):
        _lib.call("drb_create_control_points_noalloc", pc.data_ptr(), n, C.c_float(0.25), out.data_ptr(), n,
                  C.byref(p), stream_ptr(dev))
    # a grid too large (16 bytes per cell): refused after the bounding box, before the sort
    pts = np.array([[0.0, 0.0], [4000.0, 4000.0]], np.float32)
    d = t.as_tensor(pts).to(dev)
    with pytest.raises(_lib.DrbError, match="CAPACITY"):
        _lib.call("drb_create_control_points_noalloc", d.data_ptr(), 2, C.c_float(0.25), out.data_ptr(), 2,
                  C.byref(p), stream_ptr(dev))
    # the array is still usable after a refusal
    pc2 = _components(500, 3, 0.0, 6.0)
    ref, pr = _thin("stage", pc2, 0.25)
    got, pg = _thin("noalloc", pc2, 0.25)
    assert pg == pr and np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_noalloc_call_takes_no_device_memory():
    t = torch()
    dev = t.device("cuda:0")
    pc = _components(20_000, 4, -30.0, 30.0)
    d_pc = t.as_tensor(pc).to(dev)
    out = t.empty((len(pc), 2), dtype=t.float32, device=dev)
    p = C.c_int64(0)
    args = (d_pc.data_ptr(), len(pc), C.c_float(0.25), out.data_ptr(), len(pc), C.byref(p), stream_ptr(dev))
    _lib.call("drb_create_control_points_noalloc", *args)  # first call on the device: looks the array up
    t.cuda.synchronize()
    free0, _ = t.cuda.mem_get_info(dev)
    for _ in range(3):
        _lib.call("drb_create_control_points_noalloc", *args)
    t.cuda.synchronize()
    free1, _ = t.cuda.mem_get_info(dev)
    assert free1 == free0
