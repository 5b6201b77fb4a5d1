"""Diagonal Coulomb evolution drop-in (ffsim.apply_diag_coulomb_evolution).

Follows python/ffsim/gates/diag_coulomb.py:68-275.  The matrices are brought
into the form that ffsim's _get_mat_exp exponentiates (diagonals of the
same-spin matrices halved, None meaning no interaction); the GPU kernel then
computes E(a, b) per determinant and multiplies by exp(-i t E) in one pass.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd.gates.orbital_rotation import (
    conjugate_orbital_rotation,
    orbital_rotation_in_place,
)


def _halved(m, norb):
    m = np.array(m, dtype=float, copy=True)
    if m.shape != (norb, norb):
        raise ValueError(f"expected shape {(norb, norb)}, got {m.shape}")
    m[np.diag_indices(norb)] *= 0.5
    return np.ascontiguousarray(m)


def prepared_mats(mat, norb: int):
    """(aa, ab, bb) real matrices with same-spin diagonals halved; None stays None."""
    if isinstance(mat, np.ndarray) and mat.ndim == 2:
        aa = _halved(mat, norb)
        ab = np.ascontiguousarray(mat, dtype=float)
        return aa, ab, aa
    mat_aa, mat_ab, mat_bb = mat
    aa = _halved(mat_aa, norb) if mat_aa is not None else None
    bb = _halved(mat_bb, norb) if mat_bb is not None else None
    ab = _device.host_mat(mat_ab, (norb, norb), dtype=float) if mat_ab is not None else None
    return aa, ab, bb


def diag_coulomb_evolutions_in_place(
    states, mat, time: float, norb: int, nelec, z_representation: bool
) -> list:
    """One diagonal Coulomb evolution on several resident states of one shape and
    dtype, in one C-ABI call (one elementwise launch over all of them); each
    result is bitwise what :func:`diag_coulomb_evolution_in_place` gives."""
    states = list(states)
    if not states:
        return states
    aa, ab, bb = prepared_mats(mat, norb)
    na, nb = nelec
    tag = _device.dtype_tag(states[0])
    for t in states:
        _device.check_size(t, norb, nelec)
        if _device.dtype_tag(t) != tag or t.device != states[0].device:
            raise ValueError("a batch needs states of one dtype on one device")
    ptrs = (ctypes.c_void_p * len(states))(*[t.data_ptr() for t in states])
    _lib.call(
        "ffsim_apply_diag_coulomb_evolution_batched",
        len(states),
        ptrs,
        tag,
        norb,
        na,
        nb,
        _device.host_ptr(aa),
        _device.host_ptr(ab),
        _device.host_ptr(bb),
        float(time),
        1 if z_representation else 0,
        _device.stream_ptr(),
    )
    return states


def diag_coulomb_evolution_in_place(
    t: torch.Tensor, mat, time: float, norb: int, nelec, z_representation: bool
) -> torch.Tensor:
    aa, ab, bb = prepared_mats(mat, norb)
    na, nb = nelec
    _device.check_size(t, norb, nelec)
    _lib.call(
        "ffsim_apply_diag_coulomb_evolution",
        _device.ptr(t),
        _device.dtype_tag(t),
        norb,
        na,
        nb,
        _device.host_ptr(aa),
        _device.host_ptr(ab),
        _device.host_ptr(bb),
        float(time),
        1 if z_representation else 0,
        _device.stream_ptr(),
    )
    return t


def apply_diag_coulomb_evolution(
    vec,
    mat,
    time: float,
    norb: int,
    nelec,
    *,
    orbital_rotation=None,
    z_representation: bool = False,
    copy: bool = True,
):
    r"""Apply time evolution by a (rotated) diagonal Coulomb operator.

    Applies :math:`\mathcal{U} \exp(-i t \sum J^{(\sigma\tau)}_{ij} n_{i\sigma}
    n_{j\tau} / 2) \mathcal{U}^\dagger`.  Same arguments as ffsim's function;
    ``vec`` may be a numpy array or a resident torch tensor on the GPU.
    """
    if isinstance(nelec, int):
        if z_representation:
            raise NotImplementedError
        nelec = (nelec, 0)
    st = _device.to_device(vec, copy)
    t = st.tensor
    if orbital_rotation is not None:
        orbital_rotation_in_place(t, conjugate_orbital_rotation(orbital_rotation), norb, nelec)
    diag_coulomb_evolution_in_place(t, mat, time, norb, nelec, z_representation)
    if orbital_rotation is not None:
        orbital_rotation_in_place(t, orbital_rotation, norb, nelec)
    return st.restore(t)
