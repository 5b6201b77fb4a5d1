"""The library's host code and the oracle's C kernels under AddressSanitizer and
UndefinedBehaviorSanitizer (CPU only).

Two drivers are built with ``-fsanitize=address,undefined`` and run as child
processes; any out-of-bounds access, leak, use after free or undefined
behaviour aborts the driver, and each driver also checks its results against
closed forms (see the drivers' headers):

* ``tests/native/sanitize_host.cpp`` with ``ffsim_amd/csrc/host_core.cpp`` and
  ``tests/native/pass_emulator.cpp``: the host planner, window and group
  tables, wave schedules and the emulated fused passes;
* ``tests/native/sanitize_oracle.cpp`` with ``oracle/ffsim_oracle.c``: the
  oracle's Givens, phase, transpose, diagonal evolution, contraction and
  decomposition kernels, single- and multi-threaded.
"""

from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not available")


def _run(binary: Path, timeout: int = 300):
    env = dict(os.environ)
    env.update(ENV)
    res = subprocess.run([str(binary)], capture_output=True, text=True, env=env, timeout=timeout)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-4000:]
    assert "OK" in res.stdout, out[-4000:]
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out, out[-4000:]


def test_host_planner_and_emulated_passes_sanitized(tmp_path):
    _need("g++")
    exe = tmp_path / "sanitize_host"
    subprocess.run(["g++", "-std=c++17", *SAN, str(ROOT / "tests/native/sanitize_host.cpp"),
                    str(ROOT / "tests/native/pass_emulator.cpp"),
                    str(ROOT / "ffsim_amd/csrc/host_core.cpp"), "-o", str(exe)],
                   check=True, capture_output=True)
    _run(exe)


def test_oracle_kernels_sanitized(tmp_path):
    _need("g++")
    _need("gcc")
    drv, orc, exe = tmp_path / "drv.o", tmp_path / "oracle.o", tmp_path / "sanitize_oracle"
    subprocess.run(["g++", "-std=c++17", *SAN, "-c", str(ROOT / "tests/native/sanitize_oracle.cpp"),
                    "-o", str(drv)], check=True, capture_output=True)
    subprocess.run(["gcc", "-std=c11", *SAN, "-pthread", "-c", str(ROOT / "oracle/ffsim_oracle.c"),
                    "-o", str(orc)], check=True, capture_output=True)
    subprocess.run(["g++", *SAN, "-pthread", str(drv), str(orc), "-o", str(exe), "-lm"],
                   check=True, capture_output=True)
    _run(exe)
