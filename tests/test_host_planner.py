"""Host planner, fiber and group tables, checked on the CPU.

tests/native/pass_emulator.cpp runs the exact programs the library builds
for the GPU fiber kernel (window passes with transposed stores, register
sub-passes, layer pairs, group tables) with plain loops; the result must
match the oracle.  Small windows force many passes and sub-windows.
"""

from __future__ import annotations

import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("emu") / "libpass_emulator.so"
    subprocess.run(
        ["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
         str(ROOT / "tests/native/pass_emulator.cpp"),
         str(ROOT / "ffsim_amd/csrc/host_core.cpp"), "-o", str(out)],
        check=True, capture_output=True)
    lib = ctypes.CDLL(str(out))
    P = ctypes.c_void_p
    lib.emulate.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int,
                            ctypes.c_int]
    return lib


@pytest.mark.parametrize("merge", [0, 1])
@pytest.mark.parametrize("norb,nelec,w", [
    (4, (2, 2), 4), (5, (2, 3), 5), (6, (3, 2), 4), (7, (0, 3), 4), (7, (7, 2), 4),
    (8, (4, 4), 5), (8, (3, 5), 6), (9, (2, 7), 5), (10, (5, 4), 6), (12, (6, 6), 8),
    (12, (6, 6), 12), (13, (6, 5), 12)])
def test_emulated_passes_match_oracle(emulator, norb, nelec, w, merge, rng):
    """merge=1: groups with more than half the sub-window occupied are stored as
    complemented groups of the mirrored class (the c64 tables)."""
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    vec = O.random_state_vector(da * db, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    expected = O.apply_orbital_rotation(vec, (Ua, Ub), norb, nelec)
    x = vec.copy()
    work = np.zeros_like(x)
    p = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ua, ub = np.ascontiguousarray(Ua), np.ascontiguousarray(Ub)
    emulator.emulate(p(x), p(work), norb, nelec[0], nelec[1], p(ua), p(ub), w, merge)
    np.testing.assert_allclose(x, expected, atol=1e-12, rtol=0)
