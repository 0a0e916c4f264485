"""The C-ABI library builds, loads and exports every entry point include/direbm_hip.h declares.

No compute call is made (no GPU here); the Python layer must refuse to run without a GPU rather
than fall back to the CPU.
"""

import ctypes
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "direbm_hip.h"


def _declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(drb_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ["drb_collide", "drb_equilibrium", "drb_disperse", "drb_create_control_points",
              "drb_refine_control_points", "drb_resample", "drb_sim_create", "drb_sim_step", "drb_lbm_step"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    from direbm_amd import _lib

    h = _lib.lib()
    for name in _declared():
        assert hasattr(h, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (drb_\w+)", out))
    assert set(_declared()) <= exported
    assert set(_lib.SIGNATURES) <= exported


def test_version_and_error_string_without_gpu():
    from direbm_amd import _lib

    h = _lib.lib()
    assert h.drb_version() >= 100
    assert isinstance(h.drb_last_error(), bytes)


def test_invalid_arguments_are_rejected_without_touching_the_device():
    from direbm_amd import _lib

    h = _lib.lib()
    assert h.drb_collide(None, 5, 7, 1, ctypes.c_float(0.6), None) == 1
    assert b"bad array" in h.drb_last_error()
    assert h.drb_collide(None, 0, 7, 1, ctypes.c_float(-1.0), None) == 1
    assert h.drb_lbm_step(None, None, 4, 4, ctypes.c_float(0.6), 1, None) == 1
    assert h.drb_sim_create(None, None, 1, None) == 1
    assert h.drb_sim_step(None, 1) == 1


def test_no_cpu_fallback():
    import torch

    import direbm_amd as drb

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(drb.NoDeviceError):
        drb.collide(np.ones((2, 7)), 0.6)
    with pytest.raises(drb.NoDeviceError):
        drb.GpuSimulator(np.zeros((1, 2)), np.ones((1, 7)), 0.6)
    with pytest.raises(drb.NoDeviceError):
        drb.GpuHexLBM(4, 4, 0.6)


def test_product_package_does_not_import_the_oracle():
    for p in (ROOT / "direbm_amd").rglob("*.py"):
        src = p.read_text()
        assert "import oracle" not in src and "from oracle" not in src, p


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr
