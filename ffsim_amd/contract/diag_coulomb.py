"""Diagonal Coulomb contraction drop-ins (python/ffsim/contract/diag_coulomb.py)."""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd._linop import DeviceLinearOperator
from ffsim_amd.gates.orbital_rotation import (
    conjugate_orbital_rotation,
    orbital_rotation_in_place,
)


def get_mats(mat, norb: int, z_representation: bool):
    """Matrix conventions of python/ffsim/contract/diag_coulomb.py:102-130
    (number representation: same-spin diagonals halved; None -> zeros)."""
    if isinstance(mat, np.ndarray):
        aa = np.array(mat, dtype=float, copy=True)
        ab = np.ascontiguousarray(mat, dtype=float)
        if not z_representation:
            aa[np.diag_indices(norb)] *= 0.5
        return np.ascontiguousarray(aa), ab, np.ascontiguousarray(aa)
    mat_aa, mat_ab, mat_bb = mat
    out = []
    for m, halve in ((mat_aa, True), (mat_ab, False), (mat_bb, True)):
        if m is None:
            out.append(None)
            continue
        m = np.array(m, dtype=float, copy=True)
        if halve and not z_representation:
            m[np.diag_indices(norb)] *= 0.5
        out.append(np.ascontiguousarray(m))
    return tuple(out)


def contract_diag_coulomb_into(
    src: torch.Tensor,
    out: torch.Tensor,
    mat,
    norb: int,
    nelec,
    z_representation: bool = False,
    scale: float = 1.0,
    accumulate: bool = False,
) -> torch.Tensor:
    """out (=|+=) scale * E(a, b) * src on the device."""
    aa, ab, bb = get_mats(mat, norb, z_representation)
    na, nb = nelec
    _device.check_size(src, norb, nelec)
    _lib.call(
        "ffsim_contract_diag_coulomb",
        _device.ptr(src),
        _device.ptr(out),
        _device.dtype_tag(src),
        norb,
        na,
        nb,
        _device.host_ptr(aa),
        _device.host_ptr(ab),
        _device.host_ptr(bb),
        1 if z_representation else 0,
        float(scale),
        1 if accumulate else 0,
        _device.stream_ptr(),
    )
    return out


def contract_diag_coulombs_in_place(states, mats, norb: int, nelec,
                                    z_representation: bool = False, scale: float = 1.0) -> list:
    """state_i = scale * E_i(a, b) * state_i for resident states of one shape and
    dtype, each with its own diagonal Coulomb operator (``mats[i]`` in any form
    :func:`contract_diag_coulomb_into` takes; the same entries None in all), in
    one C-ABI call and one elementwise launch; each result is bitwise what
    :func:`contract_diag_coulomb_into` gives with ``out`` the state itself."""
    states, mats = list(states), list(mats)
    if len(mats) != len(states):
        raise ValueError("one operator per state")
    if not states:
        return states
    prepared = [get_mats(m, norb, z_representation) for m in mats]
    tag = _device.dtype_tag(states[0])
    for t in states:
        _device.check_size(t, norb, nelec)
        if _device.dtype_tag(t) != tag or t.device != states[0].device:
            raise ValueError("a batch needs states of one dtype on one device")
    stacked = []
    for k in range(3):
        present = [p[k] is not None for p in prepared]
        if any(present) and not all(present):
            raise ValueError("a batch needs the same interactions in every operator")
        stacked.append(np.ascontiguousarray(np.stack([p[k] for p in prepared]))
                       if present[0] else None)
    na, nb = nelec
    ptrs = (ctypes.c_void_p * len(states))(*[t.data_ptr() for t in states])
    _lib.call(
        "ffsim_contract_diag_coulomb_batched",
        len(states),
        ptrs,
        tag,
        norb,
        na,
        nb,
        _device.host_ptr(stacked[0]),
        _device.host_ptr(stacked[1]),
        _device.host_ptr(stacked[2]),
        1 if z_representation else 0,
        float(scale),
        _device.stream_ptr(),
    )
    return states


def contract_diag_coulomb(vec, mat, norb: int, nelec, *, z_representation: bool = False):
    r"""Contract a diagonal Coulomb operator with a vector.

    Returns :math:`\sum Z^{(\sigma\tau)}_{ij} n_{i\sigma} n_{j\tau} / 2 \cdot
    \text{vec}` (python/ffsim/contract/diag_coulomb.py:41-99).
    """
    st = _device.to_device(vec, copy=False)
    out = torch.empty_like(st.tensor)
    contract_diag_coulomb_into(st.tensor, out, mat, norb, nelec, z_representation)
    return st.restore(out)


def diag_coulomb_linop(
    mat, norb: int, nelec, *, orbital_rotation=None, z_representation: bool = False
) -> DeviceLinearOperator:
    r"""Linear operator of a (rotated) diagonal Coulomb operator
    (python/ffsim/contract/diag_coulomb.py:195-272)."""

    def matvec_device(v: torch.Tensor) -> torch.Tensor:
        t = v.clone()
        if orbital_rotation is not None:
            orbital_rotation_in_place(t, conjugate_orbital_rotation(orbital_rotation), norb, nelec)
        contract_diag_coulomb_into(t, t, mat, norb, nelec, z_representation)
        if orbital_rotation is not None:
            orbital_rotation_in_place(t, orbital_rotation, norb, nelec)
        return t

    return DeviceLinearOperator(norb, nelec, matvec_device)
