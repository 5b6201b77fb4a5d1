r"""One-body contraction drop-ins (python/ffsim/contract/one_body.py).

``contract_one_body`` applies :math:`\sum_{ij\sigma} M_{ij} a^\dagger_{i\sigma}
a_{j\sigma}` for a complex M, the operator ffsim builds from four real calls
of pyscf's ``contract_1e`` over cached link tables (one_body.py:24-56).  Here
it is one C-ABI call: window passes over device-resident link tables,
complex arithmetic accumulated in FP64.  For a Hermitian M on a state large
enough to need the transposition buffer it is instead
:math:`W(V) \, (\sum_p \lambda_p n_p) \, W(V)^\dagger` with
:math:`M = V \mathrm{diag}(\lambda) V^\dagger`: two orbital rotations and one
diagonal pass, which move fewer bytes than the gathers there.
"""

from __future__ import annotations

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd._linop import DeviceLinearOperator


def _eigen_path(m: np.ndarray, src: torch.Tensor, norb: int, nelec) -> bool:
    """Whether the contraction runs as rotate / number operators / rotate back:
    for an exactly Hermitian M on a state that needs the transposition buffer
    (norb >= 13), where the
    window passes' gathers cost more than two orbital rotations (measured at
    norb=16 (8,8): 43.4 against 30.5 ms in complex128, 34.7 against 18.9 ms in
    complex64)."""
    return (norb >= 2 and np.array_equal(m, m.conj().T)
            and _device.work_elements(src, norb, *nelec) > 0)


def contract_one_body_into(src: torch.Tensor, out: torch.Tensor, mat, norb: int, nelec,
                           scale=1.0, accumulate=False) -> torch.Tensor:
    """out (+)= scale * M^ src on the device; ``out`` must be a different tensor."""
    m = np.ascontiguousarray(np.asarray(mat, dtype=np.complex128))
    if m.shape != (norb, norb):
        raise ValueError(f"expected a ({norb}, {norb}) matrix, got {m.shape}")
    na, nb = nelec
    _device.check_size(src, norb, nelec)
    if _eigen_path(m, src, norb, nelec):
        from ffsim_amd.contract.num_op_sum import contract_num_op_sum_into
        from ffsim_amd.gates.orbital_rotation import orbital_rotation_in_place
        from ffsim_amd.hamiltonians.double_factorized_hamiltonian import axpby_

        eigs, vecs = np.linalg.eigh(m)
        t = src.clone()
        orbital_rotation_in_place(t, vecs.T.conj(), norb, nelec)
        contract_num_op_sum_into(t, t, eigs, norb, nelec)
        orbital_rotation_in_place(t, vecs, norb, nelec)
        if accumulate:
            axpby_(out, out, t, 1.0, scale)
        else:
            axpby_(out, t, None, scale)
        return out
    _lib.call(
        "ffsim_contract_one_body",
        _device.ptr(src),
        _device.ptr(out),
        _device.dtype_tag(src),
        norb,
        na,
        nb,
        _device.host_ptr(m),
        float(scale),
        1 if accumulate else 0,
        _device.stream_ptr(),
    )
    return out


def contract_one_body(vec, mat: np.ndarray, norb: int, nelec: tuple[int, int]):
    r"""Contract a one-body tensor with a vector (one_body.py:24-56)."""
    st = _device.to_device(vec, copy=False)
    out = torch.empty_like(st.tensor)
    contract_one_body_into(st.tensor, out, mat, norb, nelec)
    return st.restore(out)


def one_body_linop(mat: np.ndarray, norb: int, nelec: tuple[int, int]) -> DeviceLinearOperator:
    r"""Linear operator of a one-body tensor (one_body.py:59-90); ``rmatvec`` applies
    the conjugate transpose."""
    mat = np.asarray(mat, dtype=np.complex128)
    mat_h = np.ascontiguousarray(mat.T.conj())

    def matvec_device(v: torch.Tensor) -> torch.Tensor:
        return contract_one_body_into(v, torch.empty_like(v), mat, norb, nelec)

    def rmatvec_device(v: torch.Tensor) -> torch.Tensor:
        return contract_one_body_into(v, torch.empty_like(v), mat_h, norb, nelec)

    return DeviceLinearOperator(norb, nelec, matvec_device, rmatvec_device)
