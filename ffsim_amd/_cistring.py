"""Determinant-string tables from the library's host routines.

Bit-exact replacements for ffsim._cistring.make_strings / gen_occslst
(python/ffsim/_cistring.py:21-30, pyscf cistring): ascending occupation
bitmasks and the occupied-orbital lists in the same order.
"""

from __future__ import annotations

import ctypes
import functools
import math

import numpy as np

from ffsim_amd import _lib


def _range_norb(orbitals) -> int:
    orbitals = list(orbitals)
    if orbitals != list(range(len(orbitals))):
        raise ValueError("only range(norb) orbital lists are supported")
    return len(orbitals)


@functools.cache
def _strings(norb: int, nocc: int) -> np.ndarray:
    n = math.comb(norb, nocc)
    out = np.zeros(n, dtype=np.int64)
    if n:
        _lib.call("ffsim_make_strings", norb, nocc, out.ctypes.data_as(ctypes.c_void_p))
    out.setflags(write=False)
    return out


@functools.cache
def _occslst(norb: int, nocc: int) -> np.ndarray:
    n = math.comb(norb, nocc)
    out = np.zeros((n, nocc), dtype=np.uint)
    if n and nocc:
        _lib.call("ffsim_gen_occslst", norb, nocc, out.ctypes.data_as(ctypes.c_void_p))
    out.setflags(write=False)
    return out


def make_strings(orbitals, nocc: int) -> np.ndarray:
    return _strings(_range_norb(orbitals), nocc)


def gen_occslst(orbitals, nocc: int) -> np.ndarray:
    return _occslst(_range_norb(orbitals), nocc)


def zero_one_subspace_indices(norb: int, nocc: int, target_orbs: tuple[int, int]) -> np.ndarray:
    """Addresses with orbital i occupied / j empty, then their partners
    (python/ffsim/gates/orbital_rotation.py:202-212)."""
    i, j = target_orbs
    n = 2 * math.comb(norb - 2, nocc - 1) if nocc >= 1 and norb >= 2 else 0
    out = np.zeros(max(n, 1), dtype=np.uint64)
    written = _lib.call(
        "ffsim_zero_one_subspace_indices", norb, nocc, i, j, out.ctypes.data_as(ctypes.c_void_p)
    )
    return out[:written].astype(np.uint, copy=False)
