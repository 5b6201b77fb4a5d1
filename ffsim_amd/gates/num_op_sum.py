"""Number-operator-sum evolution drop-in (ffsim.apply_num_op_sum_evolution).

Follows python/ffsim/gates/num_op_sum.py:62-236; both spin sectors are
applied by one elementwise GPU pass.
"""

from __future__ import annotations

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd.gates.orbital_rotation import (
    conjugate_orbital_rotation,
    orbital_rotation_in_place,
)


def _coeffs(c, norb):
    if c is None:
        return None
    a = np.ascontiguousarray(np.asarray(c, dtype=float).reshape(-1))
    if a.shape != (norb,):
        raise ValueError(f"expected {norb} coefficients, got {a.shape}")
    return a


def num_op_sum_evolution_in_place(t: torch.Tensor, coeffs, time: float, norb: int, nelec):
    if isinstance(coeffs, np.ndarray):
        ca = cb = _coeffs(coeffs, norb)
    else:
        ca, cb = (_coeffs(c, norb) for c in coeffs)
    na, nb = nelec
    _device.check_size(t, norb, nelec)
    _lib.call(
        "ffsim_apply_num_op_sum_evolution",
        _device.ptr(t),
        _device.dtype_tag(t),
        norb,
        na,
        nb,
        _device.host_ptr(ca),
        _device.host_ptr(cb),
        float(time),
        _device.stream_ptr(),
    )
    return t


def apply_num_op_sum_evolution(
    vec, coeffs, time: float, norb: int, nelec, *, orbital_rotation=None, copy: bool = True
):
    r"""Apply time evolution by a (rotated) linear combination of number operators.

    Applies :math:`\mathcal{U} \exp(-i t \sum_{i\sigma} \lambda^{(\sigma)}_i
    n_{i\sigma}) \mathcal{U}^\dagger`.  Same arguments as ffsim's function.
    """
    st = _device.to_device(vec, copy)
    t = st.tensor
    if isinstance(nelec, int):
        spin_nelec = (nelec, 0)
        coeffs = (coeffs, None)
        rot = orbital_rotation
    else:
        spin_nelec = nelec
        rot = orbital_rotation
    if rot is not None:
        orbital_rotation_in_place(t, conjugate_orbital_rotation(rot), norb, nelec)
    num_op_sum_evolution_in_place(t, coeffs, time, norb, spin_nelec)
    if rot is not None:
        orbital_rotation_in_place(t, rot, norb, nelec)
    return st.restore(t)
