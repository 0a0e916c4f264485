"""Build libdirebm_hip.so (gfx950) in-tree with hipcc.

Usage: python -m direbm_amd.build   (or direbm_amd.build.build() from Python).

Flags: -O3, -ffp-contract=off (one rounding per fp32 operation, needed for bit-exact cell keys
and radius predicates), HIP's default IEEE fp32 division / square root, and packed-fp32
instructions disabled in the device code.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libdirebm_hip.so"
SOURCES = ["sort.hip", "physics.hip", "step.hip", "adjoint.hip", "replay.hip", "fields.hip", "step3.hip", "lbm.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DIREBM_ARCH", "gfx950")

FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False, defines=(), out: Path | None = None) -> Path:
    """Compile every source to an object and link the shared library (always rebuilds)."""
    lib = LIB if out is None else Path(out)
    objdir = HERE / "_build" / (lib.stem if out is not None else "default")
    objdir.mkdir(parents=True, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(str(obj))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")
    tmp = lib.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    print(p)
