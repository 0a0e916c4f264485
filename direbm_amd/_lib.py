"""ctypes binding of libdirebm_hip.so (the C-ABI declared in include/direbm_hip.h).

Loading the library needs no GPU; every compute entry point needs one and the Python layer
raises ``NoDeviceError`` before calling it when no HIP device is visible.  There is no CPU
fallback: if the library is missing, ``lib()`` raises ImportError.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().with_name("libdirebm_hip.so")

DRB_OK = 0
ERRORS = {
    1: "INVALID_ARG",
    2: "CAPACITY",
    3: "KEY_RANGE",
    4: "NONFINITE",
    5: "HIP",
    6: "STATE",
}


class DrbError(RuntimeError):
    """A non-OK status from the engine; .code holds the DRB_ERR_* value."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"DRB_ERR_{ERRORS.get(code, code)}: {msg}")
        self.code = code


class NoDeviceError(RuntimeError):
    """No HIP GPU is visible: the engine has no CPU path."""


class Params(C.Structure):
    """drb_params (include/direbm_hip.h)."""

    _fields_ = [
        ("tau", C.c_float),
        ("dx", C.c_float),
        ("alpha", C.c_float),
        ("kappa_hard", C.c_int32),
        ("rho_rest", C.c_float),
        ("r_res", C.c_float),
        ("eps", C.c_float),
        ("periodic", C.c_int32),
        ("box", C.c_float),
        ("two_level", C.c_int32),
        ("cx", C.c_float),
        ("cy", C.c_float),
        ("r_in", C.c_float),
        ("kappa_soft", C.c_int32),
    ]


_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_F = C.c_float

# name -> (argtypes)
SIGNATURES = {
    "drb_last_error": [],
    "drb_version": [],
    "drb_collide": [_P, _I64, _I64, _I64, _F, _P],
    "drb_equilibrium": [_P, _P, _I64, _P, _I64, _I64, _P],
    "drb_disperse": [_P, _P, _I64, _F, _P, _P, _P, _P],
    "drb_stage_create": [_P],
    "drb_stage_destroy": [_P],
    "drb_create_control_points": [_P, _P, _I64, _F, _P, _I64, _P, _P],
    "drb_create_control_points_noalloc": [_P, _I64, _F, _P, _I64, _P, _P],
    "drb_radix_sort_pairs": [_P, _P, _P, _I64, _I32, _P],
    "drb_refine_control_points": [_P, _P, _I64, _P, _P, _P, _I64, _F, C.c_int, _P, _P],
    "drb_resample": [_P, _P, _P, _P, _I64, _P, _I64, _F, _P, _P, _P, _P],
    "drb_bin_fields": [_P, _P, _P, _I64, _P, _I32, _P, _P, _P],
    "drb_bin_fields_vjp": [_P, _I64, _P, _I32, _P, _P, _P, _P],
    "drb_sim_create": [_P, C.POINTER(Params), _I32, _P],
    "drb_sim_destroy": [_P],
    "drb_sim_set_state": [_P, _P, _P, _I64, _P],
    "drb_sim_step": [_P, _I32],
    "drb_sim_set_graph": [_P, _I32],
    "drb_sim_set_graph_horizon": [_P, _I32],
    "drb_sim_graph_info": [_P, _P],
    "drb_sim_set_obstacle": [_P, _I32, _F, _F, _F],
    "drb_sim_autotune": [_P, _P],
    "drb_sim_set_lanes": [_P, _I32, _I32],
    "drb_sim_set_tiles": [_P, _I32, _I32, _I32, _I32],
    "drb_sim_get_tiles": [_P, _P],
    "drb_sim_count": [_P, _P, _P],
    "drb_sim_get_state": [_P, _P, _P],
    "drb_sim_diagnostics": [_P, _P],
    "drb_sim_set_timing": [_P, _I32],
    "drb_sim_stage_ms": [_P, _P],
    "drb_sim_updates": [_P, _P],
    "drb_collide_vjp": [_P, _I64, _I64, _I64, _F, _P, _P, _P, _P],
    "drb_equilibrium_vjp": [_P, _P, _I64, _P, _I64, _I64, _P, _P, _P],
    "drb_sim_set_state_device": [_P, _P, _P, _I64, _P],
    "drb_sim_get_state_device": [_P, _P, _P],
    "drb_sim_tape": [_P, _I32],
    "drb_sim_tape_info": [_P, _I32, _P, _P],
    "drb_sim_step_vjp": [_P, _I32, _P, _P, _P, _P, _P],
    "drb_sim_tape_vjp": [_P, _P, _P, _P, _P, _P],
    "drb_sim_replay": [_P, _I32, _F, _P, _P, _P, _P],
    "drb_sim_replay64": [_P, _I32, C.c_double, _P, _P, _P, _P],
    "drb_d3q13": [_P, _P],
    "drb_sim3_create": [_P, C.POINTER(Params), _P],
    "drb_sim3_destroy": [_P],
    "drb_sim3_set_sphere": [_P, _F, _F, _F, _F],
    "drb_sim3_set_state": [_P, _P, _P, _I64],
    "drb_sim3_step": [_P, _I32],
    "drb_sim3_set_lanes": [_P, _I32, _I32],
    "drb_sim3_count": [_P, _P],
    "drb_sim3_get_state": [_P, _P, _P],
    "drb_sim3_diagnostics": [_P, _P],
    "drb_lbm_step": [_P, _P, _I32, _I32, _F, _I32, _P],
    "drb_lbm_set_node": [_P, _I32, _I32, _I32, _I32, _F, _P],
    "drb_lbm_init": [_P, _P, _I32, _I32, _P],
}

_lib = None


def lib():
    """The loaded library (ImportError if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m direbm_amd.build`")
        # PyTorch ships its own HIP runtime.  Load it first so that the library binds to that same
        # runtime: loaded the other way round, the process holds two runtimes and the library's
        # finds no device.
        import torch  # noqa: F401

        h = C.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name == "drb_last_error" else C.c_int
        _lib = h
    return _lib


def check(code: int):
    if code != DRB_OK:
        raise DrbError(code, (lib().drb_last_error() or b"").decode())


def call(name: str, *args):
    check(getattr(lib(), name)(*args))
