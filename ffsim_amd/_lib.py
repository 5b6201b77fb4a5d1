"""ctypes binding of the ffsim_amd C ABI (include/ffsim_amd.h)."""

from __future__ import annotations

import ctypes
import functools
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libffsim_amd.so"

C128 = 0
C64 = 1

_ERR_VALUE = -1
_ERR_ASSERT = -2
_ERR_HIP = -3
_ERR_NOTIMPL = -4

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_I64 = ctypes.c_int64

# name -> argtypes, in the order of include/ffsim_amd.h
SIGNATURES: dict[str, list] = {
    "ffsim_amd_last_error": [],
    "ffsim_amd_version": [],
    "ffsim_amd_init": [_I],
    "ffsim_amd_profile_begin": [],
    "ffsim_amd_profile_end": [_P],
    "ffsim_make_strings": [_I, _I, _P],
    "ffsim_gen_occslst": [_I, _I, _P],
    "ffsim_zero_one_subspace_indices": [_I, _I, _I, _I, _P],
    "ffsim_givens_decomposition": [_P, _I, _D, _P, _P, _P, _P, _P, _P],
    "ffsim_orbital_rotation_work_elements": [_I, _I, _I, _I],
    "ffsim_apply_orbital_rotation": [_P, _P, _I, _I, _I, _I, _P, _P, _P],
    "ffsim_apply_orbital_rotation_batched": [_I, _P, _P, _I, _I, _I, _I, _P, _P, _P],
    "ffsim_apply_diag_coulomb_evolution": [_P, _I, _I, _I, _I, _P, _P, _P, _D, _I, _P],
    "ffsim_apply_diag_coulomb_evolution_batched": [_I, _P, _I, _I, _I, _I, _P, _P, _P, _D, _I,
                                                   _P],
    "ffsim_apply_num_op_sum_evolution": [_P, _I, _I, _I, _I, _P, _P, _D, _P],
    "ffsim_contract_diag_coulomb": [_P, _P, _I, _I, _I, _I, _P, _P, _P, _I, _D, _I, _P],
    "ffsim_contract_diag_coulomb_batched": [_I, _P, _I, _I, _I, _I, _P, _P, _P, _I, _D, _P],
    "ffsim_contract_num_op_sum": [_P, _P, _I, _I, _I, _I, _P, _P, _D, _I, _P],
    "ffsim_axpby": [_P, _P, _P, _I, _I64, _D, _D, _D, _D, _P],
    "ffsim_apply_ucj_spin_balanced": [_P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P],
    "ffsim_apply_ucj_spin_balanced_batched": [_I, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _P],
    "ffsim_contract_one_body": [_P, _P, _I, _I, _I, _I, _P, _D, _I, _P],
    "ffsim_diag_coulomb_grad_reduce": [_P, _P, _I, _I, _I, _I, _P, _P],
    "ffsim_transition_rdm": [_P, _P, _I, _I, _I, _I, _P, _P],
    "ffsim_apply_diag_coulomb_sequence": [_P, _P, _I, _I, _I, _I, _I, _P, _P, _P, _I, _P],
    "ffsim_apply_givens_rotation_in_place": [_P, _I, _I64, _I64, _D, _D, _D, _P, _P, _I64, _P],
    "ffsim_apply_phase_shift_in_place": [_P, _I, _I64, _I64, _D, _D, _P, _I64, _P],
    "ffsim_apply_givens_layer": [_P, _I, _I, _I, _I64, _I, _I, _P, _P, _P, _P, _P],
    "ffsim_apply_diag_coulomb_evolution_noalloc": [_P, _I, _I, _I, _I, _P, _P, _P, _D, _I, _P],
}


class LibraryMissingError(RuntimeError):
    """Raised when the HIP library has not been built."""


@functools.cache
def lib() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise LibraryMissingError(
            f"{LIB_PATH} is missing: build it with `python -m ffsim_amd._build` "
            "(there is no CPU fallback for the gate kernels)"
        )
    handle = ctypes.CDLL(str(LIB_PATH))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.argtypes = argtypes
        if name == "ffsim_amd_last_error":
            fn.restype = ctypes.c_char_p
        elif name == "ffsim_orbital_rotation_work_elements":
            fn.restype = ctypes.c_int64
        else:
            fn.restype = ctypes.c_int
    return handle


def check(status: int) -> int:
    """Map a C status to the Python exception the reference raises."""
    if status >= 0:
        return status
    msg = lib().ffsim_amd_last_error().decode(errors="replace")
    if status == _ERR_VALUE:
        raise ValueError(msg)
    if status == _ERR_ASSERT:
        raise AssertionError(msg)
    if status == _ERR_NOTIMPL:
        raise NotImplementedError(msg)
    raise RuntimeError(f"HIP error: {msg}")


def call(name: str, *args) -> int:
    return check(getattr(lib(), name)(*args))
