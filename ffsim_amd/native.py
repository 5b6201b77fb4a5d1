"""ffsim's native per-rotation calls on device-resident states.

ffsim's Python schedule (python/ffsim/gates/orbital_rotation.py:117-154)
calls its compiled module once per Givens rotation and once per orbital
phase: ``_lib.apply_givens_rotation_in_place(vec, c, s, slice1, slice2)`` and
``_lib.apply_phase_shift_in_place(vec, phase, indices)``
(python/ffsim/_lib.pyi:36-41, src/gates/orbital_rotation.rs:19-104,
src/gates/phase_shift.rs:16-30).  The functions here take the same arguments,
with ``vec`` a 2-D torch tensor on the GPU (complex128 or complex64), updated
in place on the current stream; the index arrays are ffsim's host tables.

:func:`apply_givens_layer` applies a whole layer of disjoint adjacent
rotations of one spin sector in one launch, with the pair tables built and
cached on the device.

These are the fine-grained form of the boundary.  The fused entry points
(``apply_orbital_rotation`` and the rest) remain the fast path: one HBM pass
there applies dozens of rotations.
"""

from __future__ import annotations

import numpy as np
import torch

from ffsim_amd import _device, _lib


def _matrix(vec: torch.Tensor) -> tuple[int, int]:
    if not isinstance(vec, torch.Tensor) or vec.device.type != "cuda":
        raise TypeError("vec must be a torch tensor on the GPU")
    if vec.ndim != 2 or not vec.is_contiguous():
        raise ValueError("vec must be a C-contiguous 2-D tensor")
    return vec.shape[0], vec.shape[1]


def _indices(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a), dtype=np.uint64)


def apply_givens_rotation_in_place(vec: torch.Tensor, c: float, s: complex, slice1,
                                   slice2) -> None:
    """Rows ``slice1[k]``, ``slice2[k]`` become ``c x + s y`` and ``c y - conj(s) x``."""
    n_rows, n_cols = _matrix(vec)
    i1, i2 = _indices(slice1), _indices(slice2)
    if i1.shape != i2.shape or i1.ndim != 1:
        raise ValueError("slice1 and slice2 must be 1-D arrays of the same length")
    s = complex(s)
    _lib.call("ffsim_apply_givens_rotation_in_place", _device.ptr(vec),
              _device.dtype_tag(vec), n_rows, n_cols, float(c), s.real, s.imag,
              _device.host_ptr(i1), _device.host_ptr(i2), len(i1), _device.stream_ptr())


def apply_phase_shift_in_place(vec: torch.Tensor, phase: complex, indices) -> None:
    """Rows ``indices[k]`` of ``vec`` are multiplied by ``phase``."""
    n_rows, n_cols = _matrix(vec)
    idx = _indices(indices)
    if idx.ndim != 1:
        raise ValueError("indices must be a 1-D array")
    phase = complex(phase)
    _lib.call("ffsim_apply_phase_shift_in_place", _device.ptr(vec), _device.dtype_tag(vec),
              n_rows, n_cols, phase.real, phase.imag, _device.host_ptr(idx), len(idx),
              _device.stream_ptr())


def apply_givens_layer(vec: torch.Tensor, rotations, norb: int, nocc: int, *,
                       strings_on_rows: bool = True) -> None:
    """Apply disjoint adjacent rotations ``(c, s, i, j)`` to one spin sector in one pass.

    Each rotation acts as ``apply_givens_rotation_in_place(vec, c, s, slice1,
    slice2)`` with ``slice1`` the strings with ``i`` occupied and ``j`` empty and
    ``slice2`` their partners (python/ffsim/gates/orbital_rotation.py:202-212).
    With ``strings_on_rows`` the sector's strings index the rows of ``vec``;
    otherwise its columns (the beta sector of a ``dim_a x dim_b`` state, rotated
    in place without a transpose).
    """
    n_rows, n_cols = _matrix(vec)
    rots = list(rotations)
    n_other = n_cols if strings_on_rows else n_rows
    c = np.array([float(r[0]) for r in rots], dtype=np.float64)
    s = np.array([complex(r[1]) for r in rots], dtype=np.complex128)
    i = np.array([int(r[2]) for r in rots], dtype=np.int32)
    j = np.array([int(r[3]) for r in rots], dtype=np.int32)
    _lib.call("ffsim_apply_givens_layer", _device.ptr(vec), _device.dtype_tag(vec), norb, nocc,
              n_other, 1 if strings_on_rows else 0, len(rots), _device.host_ptr(c),
              _device.host_ptr(s), _device.host_ptr(i), _device.host_ptr(j),
              _device.stream_ptr())
