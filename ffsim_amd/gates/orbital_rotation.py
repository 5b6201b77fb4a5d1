"""Orbital rotation drop-in (ffsim.apply_orbital_rotation).

Follows python/ffsim/gates/orbital_rotation.py:44-154: the same signature,
the same `mat` forms (one matrix, an (alpha, beta) pair with None entries, or
a stacked array) and the same copy semantics.  The work is one C-ABI call
that decomposes each unitary on the host (FP64) and applies the Givens
network in fused window passes on the GPU.
"""

from __future__ import annotations

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
