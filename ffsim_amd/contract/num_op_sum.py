"""Number-operator-sum contraction drop-ins (python/ffsim/contract/num_op_sum.py)."""

from __future__ import annotations

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd._linop import DeviceLinearOperator
from ffsim_amd.gates.orbital_rotation import orbital_rotation_in_place


def contract_num_op_sum_into(
    src: torch.Tensor, out: torch.Tensor, coeffs, norb: int, nelec, scale=1.0, accumulate=False
) -> torch.Tensor:
    c = np.ascontiguousarray(np.asarray(coeffs, dtype=float).reshape(-1))
    if c.shape != (norb,):
        raise ValueError(f"expected {norb} coefficients, got {c.shape}")
    na, nb = nelec
    _device.check_size(src, norb, nelec)
    _lib.call(
        "ffsim_contract_num_op_sum",
        _device.ptr(src),
        _device.ptr(out),
        _device.dtype_tag(src),
        norb,
        na,
        nb,
        _device.host_ptr(c),
        _device.host_ptr(c),
        float(scale),
        1 if accumulate else 0,
        _device.stream_ptr(),
    )
    return out


def contract_num_op_sum(vec, coeffs, norb: int, nelec):
    r"""Contract :math:`\sum_{i\sigma} \lambda_i n_{i\sigma}` with a vector
    (python/ffsim/contract/num_op_sum.py:27-72)."""
    st = _device.to_device(vec, copy=False)
    out = torch.empty_like(st.tensor)
    contract_num_op_sum_into(st.tensor, out, coeffs, norb, nelec)
    return st.restore(out)


def num_op_sum_linop(coeffs, norb: int, nelec, *, orbital_rotation=None) -> DeviceLinearOperator:
    r"""Linear operator of a (rotated) number-operator sum
    (python/ffsim/contract/num_op_sum.py:75-131)."""

    def matvec_device(v: torch.Tensor) -> torch.Tensor:
        t = v.clone()
        if orbital_rotation is not None:
            orbital_rotation_in_place(t, orbital_rotation.T.conj(), norb, nelec)
        contract_num_op_sum_into(t, t, coeffs, norb, nelec)
        if orbital_rotation is not None:
            orbital_rotation_in_place(t, orbital_rotation, norb, nelec)
        return t

    return DeviceLinearOperator(norb, nelec, matvec_device)
