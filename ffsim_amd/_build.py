"""Build recipe for the HIP library (gfx950), compiled in-tree."""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libffsim_amd.so"

SOURCES = ["ffsim_amd.hip", "noalloc.hip", "host_core.cpp"]
HEADERS = ["kernels.hpp", "host_core.hpp", "status.hpp", "diag_params.hpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found; the ffsim_amd HIP library cannot be built")


def build(verbose: bool = False) -> Path:
    """Compile libffsim_amd.so for gfx950 (always recompiles)."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [
        hipcc(),
        "-O3",
        "-std=c++17",
        "--offload-arch=gfx950",
        "-fPIC",
        "-shared",
        # keep the compiler from forming packed FP32 instructions
        "-Xclang",
        "-target-feature",
        "-Xclang",
        "-packed-fp32-ops",
        "-Wall",
        "-Wno-unused-function",
        *[str(CSRC / s) for s in SOURCES],
        "-L/opt/rocm/lib",
        "-lrocblas",
        "-Wl,-rpath,/opt/rocm/lib",
        "-o",
        str(tmp),
    ]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + res.stdout + res.stderr)
    if verbose:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
