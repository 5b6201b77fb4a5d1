"""Orbital rotation drop-in (ffsim.apply_orbital_rotation).

Follows python/ffsim/gates/orbital_rotation.py:44-154: the same signature,
the same `mat` forms (one matrix, an (alpha, beta) pair with None entries, or
a stacked array) and the same copy semantics.  The work is one C-ABI call
that decomposes each unitary on the host (FP64) and applies the Givens
network in fused window passes on the GPU.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from ffsim_amd import _device, _lib


def _split_mat(mat, norb: int):
    if isinstance(mat, np.ndarray) and mat.ndim == 2:
        m = _device.host_mat(mat, (norb, norb))
        return m, m
    mat_a, mat_b = mat
    return (
        _device.host_mat(mat_a, (norb, norb)),
        _device.host_mat(mat_b, (norb, norb)),
    )


def orbital_rotation_in_place(t: torch.Tensor, mat, norb: int, nelec) -> torch.Tensor:
    """Rotate a resident 1-D device state in place (no copy)."""
    if isinstance(nelec, int):
        mat_a = _device.host_mat(mat, (norb, norb))
        mat_b = None
        na, nb = nelec, 0
    else:
        mat_a, mat_b = _split_mat(mat, norb)
        na, nb = nelec
    _device.check_size(t, norb, (na, nb))
    work = _device.work_buffer(t, norb, na, nb) if mat_b is not None else None
    _lib.call(
        "ffsim_apply_orbital_rotation",
        _device.ptr(t),
        _device.ptr(work),
        _device.dtype_tag(t),
        norb,
        na,
        nb,
        _device.host_ptr(mat_a),
        _device.host_ptr(mat_b),
        _device.stream_ptr(),
    )
    return t


def orbital_rotations_in_place(states, mats, norb: int, nelec: tuple) -> list:
    """Rotate resident 1-D device states of one shape and dtype in place, state
    ``i`` by ``mats[i]`` (one matrix, or an ``(alpha, beta)`` pair whose ``None``
    entries must be the same for every state), in one C-ABI call.  Small
    states share their launches; each result is bitwise what
    :func:`orbital_rotation_in_place` gives."""
    states, mats = list(states), list(mats)
    if len(mats) != len(states):
        raise ValueError("one rotation per state")
    if not states:
        return states
    na, nb = nelec
    pairs = [_split_mat(m, norb) for m in mats]
    has_a, has_b = pairs[0][0] is not None, pairs[0][1] is not None
    tag = _device.dtype_tag(states[0])
    for (a, b), t in zip(pairs, states):
        if (a is not None) != has_a or (b is not None) != has_b:
            raise ValueError("a batch needs the same sectors rotated in every state")
        _device.check_size(t, norb, (na, nb))
        if _device.dtype_tag(t) != tag or t.device != states[0].device:
            raise ValueError("a batch needs states of one dtype on one device")
    mats_a = np.ascontiguousarray(np.stack([a for a, _ in pairs])) if has_a else None
    mats_b = np.ascontiguousarray(np.stack([b for _, b in pairs])) if has_b else None
    # the states that need a work buffer run one after another: they share one
    work = _device.work_buffer(states[0], norb, na, nb) if has_b else None
    vec_ptrs = (ctypes.c_void_p * len(states))(*[t.data_ptr() for t in states])
    work_ptrs = (ctypes.c_void_p * len(states))(
        *[work.data_ptr() if work is not None else None for _ in states])
    _lib.call("ffsim_apply_orbital_rotation_batched", len(states), vec_ptrs, work_ptrs, tag,
              norb, na, nb, _device.host_ptr(mats_a), _device.host_ptr(mats_b),
              _device.stream_ptr())
    return states


def apply_orbital_rotation(vec, mat, norb: int, nelec, *, copy: bool = True):
    r"""Apply an orbital rotation to a vector.

    Maps :math:`a^\dagger_{i\sigma} \mapsto \sum_j U^{(\sigma)}_{ji} a^\dagger_{j\sigma}`.

    Args:
        vec: The state vector (numpy array, or a torch tensor on the GPU which
            stays resident).
        mat: One unitary for both spin sectors, a pair ``(mat_a, mat_b)`` with
            ``None`` meaning no rotation of that sector, or a stacked array.
        norb: The number of spatial orbitals.
        nelec: ``int`` (spinless) or ``(n_alpha, n_beta)``.
        copy: With ``True`` the input is never modified.  With ``False`` a
            device tensor is rotated in place and returned.

    Returns:
        The rotated vector, of the same kind as ``vec``.
    """
    st = _device.to_device(vec, copy)
    orbital_rotation_in_place(st.tensor, mat, norb, nelec)
    return st.restore(st.tensor)


def conjugate_orbital_rotation(orbital_rotation):
    """Hermitian conjugate of one rotation or of an (alpha, beta) pair."""
    if isinstance(orbital_rotation, np.ndarray) and orbital_rotation.ndim == 2:
        return orbital_rotation.T.conj()
    a, b = orbital_rotation
    return (
        a.T.conj() if a is not None else None,
        b.T.conj() if b is not None else None,
    )
