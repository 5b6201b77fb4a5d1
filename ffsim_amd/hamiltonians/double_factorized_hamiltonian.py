"""Double-factorized Hamiltonian linear operator
(python/ffsim/hamiltonians/double_factorized_hamiltonian.py)."""

from __future__ import annotations

import dataclasses

import numpy as np
import scipy.linalg
import torch

from ffsim_amd import _device, _lib
from ffsim_amd import linalg as _linalg
from ffsim_amd._linop import DeviceLinearOperator
from ffsim_amd.contract.diag_coulomb import (
    contract_diag_coulomb_into,
    contract_diag_coulombs_in_place,
)
from ffsim_amd.contract.num_op_sum import contract_num_op_sum_into
from ffsim_amd.gates.orbital_rotation import orbital_rotation_in_place, orbital_rotations_in_place


def axpby_(out: torch.Tensor, x: torch.Tensor, y: torch.Tensor | None, alpha, beta=0.0):
    """out = alpha x + beta y on the device (out may alias x or y)."""
    a = complex(alpha)
    b = complex(beta)
    _lib.call(
        "ffsim_axpby",
        _device.ptr(out),
        _device.ptr(x),
        _device.ptr(y),
        _device.dtype_tag(out),
        out.numel(),
        a.real,
        a.imag,
        b.real,
        b.imag,
        _device.stream_ptr(),
    )
    return out


def _batch_terms(v: torch.Tensor, norb: int, nelec: tuple) -> bool:
    """Whether the DF matvec rotates its terms' copies of v in shared launches:
    for states without a transposition buffer (small ones), where one copy per
    term costs little memory."""
    return _device.work_elements(v, norb, *nelec) == 0


@dataclasses.dataclass(frozen=True)
class DoubleFactorizedHamiltonian:
    r"""H = sum_{pq,sigma} kappa_pq a^dag a + 1/2 sum_t sum J^t_ij n^t_i n^t_j + constant."""

    one_body_tensor: np.ndarray
    diag_coulomb_mats: np.ndarray
    orbital_rotations: np.ndarray
    constant: float = 0.0
    z_representation: bool = False

    @property
    def norb(self) -> int:
        return self.one_body_tensor.shape[0]

    @staticmethod
    def from_molecular_hamiltonian(hamiltonian, *, z_representation: bool = False,
                                   tol: float = 1e-8, max_vecs: int | None = None,
                                   optimize: bool = False, cholesky: bool = True,
                                   **optimize_kwargs) -> "DoubleFactorizedHamiltonian":
        r"""Double-factorized form of a molecular Hamiltonian
        (double_factorized_hamiltonian.py:140-222).  ``hamiltonian`` is any object
        with ``one_body_tensor``, a real ``two_body_tensor`` in chemists' order
        :math:`(pq|rs)` and ``constant`` (ffsim's MolecularHamiltonian is one).
        The one-body tensor becomes :math:`h_{pq} - \tfrac12 \sum_r (pr|rq)`
        and the two-body tensor is factored by
        :func:`ffsim_amd.linalg.double_factorized`."""
        two_body = np.asarray(hamiltonian.two_body_tensor)
        one_body = np.asarray(hamiltonian.one_body_tensor) - 0.5 * np.einsum("prrq", two_body)
        mats, rots = _linalg.double_factorized(two_body, tol=tol, max_vecs=max_vecs,
                                               optimize=optimize, cholesky=cholesky,
                                               **optimize_kwargs)
        ham = DoubleFactorizedHamiltonian(one_body_tensor=one_body, diag_coulomb_mats=mats,
                                          orbital_rotations=rots,
                                          constant=float(hamiltonian.constant))
        return ham.to_z_representation() if z_representation else ham

    def _z_correction(self):
        r"""One-body and constant terms that move the operator between the
        number and Z representations (double_factorized_hamiltonian.py:308-330):
        :math:`\sum_t U_t\,\mathrm{diag}((r_t + c_t)/2)\,U_t^\dagger`, with
        :math:`r_t, c_t` the row and column sums of :math:`J_t`, and
        :math:`\sum_t (\mathrm{tr} J_t / 4 - \sum_{ij} J_{t,ij} / 2)`."""
        mats = np.asarray(self.diag_coulomb_mats)
        rots = np.asarray(self.orbital_rotations)
        norb = self.norb
        one_body = np.zeros((norb, norb), dtype=complex)
        for j, u in zip(mats, rots):
            w = 0.5 * (j.sum(axis=1) + j.sum(axis=0))
            one_body += (u * w) @ u.conj().T
        constant = sum(0.25 * np.trace(j) - 0.5 * j.sum() for j in mats)
        return one_body, float(constant)

    def to_z_representation(self) -> "DoubleFactorizedHamiltonian":
        """The same operator in the Z representation
        (double_factorized_hamiltonian.py:107-121)."""
        if self.z_representation:
            return self
        one_body, constant = self._z_correction()
        return dataclasses.replace(self, one_body_tensor=self.one_body_tensor + one_body,
                                   constant=self.constant + constant, z_representation=True)

    def to_number_representation(self) -> "DoubleFactorizedHamiltonian":
        """The same operator in the number representation
        (double_factorized_hamiltonian.py:123-138)."""
        if not self.z_representation:
            return self
        one_body, constant = self._z_correction()
        return dataclasses.replace(self, one_body_tensor=self.one_body_tensor - one_body,
                                   constant=self.constant - constant, z_representation=False)

    def _diag_(self, norb: int, nelec) -> np.ndarray:
        r"""Diagonal entries in the determinant basis (ffsim computes them
        through the molecular Hamiltonian, double_factorized_hamiltonian.py:277-280).

        Host numpy, by Wick's theorem on each determinant ``D``: with
        :math:`\tilde n_i = \sum_{pq} U_{pi} U^*_{qi} a^\dagger_p a_q` and
        :math:`M = U^\dagger P_D U` per spin, :math:`\langle \tilde n_i\rangle = M_{ii}`,
        :math:`\langle \tilde n_i \tilde n_j\rangle = M_{ii} M_{jj} + \delta_{ij} M_{ii} - |M_{ij}|^2`
        within a spin and :math:`M^\alpha_{ii} M^\beta_{jj}` across spins."""
        from ffsim_amd.hamiltonians.diagonal_coulomb_hamiltonian import _occupation_matrix

        assert isinstance(nelec, tuple)
        occ = [_occupation_matrix(norb, n) for n in nelec]
        kappa = np.diag(self.one_body_tensor).real
        result = self.constant + (occ[0] @ kappa)[:, None] + (occ[1] @ kappa)[None, :]
        for jmat, u in zip(np.asarray(self.diag_coulomb_mats, dtype=float),
                           np.asarray(self.orbital_rotations)):
            m, same = [], []
            for o in occ:
                mm = np.einsum("pi,ap,pj->aij", u.conj(), o, u)  # M per string
                mi = np.einsum("aii->ai", mm).real
                nn = mi[:, :, None] * mi[:, None, :] - np.abs(mm) ** 2
                if self.z_representation:
                    zz = 1 - 2 * mi[:, :, None] - 2 * mi[:, None, :] + 4 * nn
                    off = jmat - np.diag(np.diag(jmat))
                    same.append(0.125 * np.einsum("ij,aij->a", off, zz))
                else:
                    nn += np.einsum("ai,ij->aij", mi, np.eye(norb))
                    same.append(0.5 * np.einsum("ij,aij->a", jmat, nn))
                m.append(mi)
            ma, mb = m
            if self.z_representation:
                za, zb = 1 - 2 * ma, 1 - 2 * mb
                cross = 0.125 * (za @ jmat @ zb.T + (zb @ jmat @ za.T).T)
            else:
                cross = 0.5 * (ma @ jmat @ mb.T + (mb @ jmat @ ma.T).T)
            result = result + same[0][:, None] + same[1][None, :] + cross
        return np.asarray(result, dtype=float).reshape(-1)

    def _approx_eq_(self, other, rtol: float, atol: float) -> bool:
        """double_factorized_hamiltonian.py:286-305."""
        if not isinstance(other, DoubleFactorizedHamiltonian):
            return NotImplemented
        return (self.z_representation == other.z_representation
                and all(np.allclose(x, y, rtol=rtol, atol=atol) for x, y in (
                    (self.constant, other.constant),
                    (self.one_body_tensor, other.one_body_tensor),
                    (self.diag_coulomb_mats, other.diag_coulomb_mats),
                    (self.orbital_rotations, other.orbital_rotations))))

    def _linear_operator_(self, norb: int, nelec) -> DeviceLinearOperator:
        """Same operator as ffsim's (double_factorized_hamiltonian.py:244-275): the
        one-body term as a rotated number-operator sum, plus one rotated
        diagonal Coulomb contraction per factor, evaluated on the GPU."""
        assert isinstance(nelec, tuple)
        eigs, vecs = scipy.linalg.eigh(self.one_body_tensor)
        mats = [np.asarray(m, dtype=float) for m in self.diag_coulomb_mats]
        rots = [np.asarray(u, dtype=np.complex128) for u in self.orbital_rotations]
        constant = self.constant
        z_rep = self.z_representation

        def matvec_device(v: torch.Tensor) -> torch.Tensor:
            result = torch.empty_like(v)
            axpby_(result, v, None, constant)
            if _batch_terms(v, norb, nelec):
                # small states (no transposition buffer): one copy per term; the
                # terms' rotations and contractions share their launches (the
                # same arithmetic per term as the loop below, summed in the
                # same order)
                ts = [v.clone() for _ in range(1 + len(mats))]
                orbital_rotations_in_place(
                    ts, [vecs.T.conj()] + [rot.T.conj() for rot in rots], norb, nelec)
                contract_num_op_sum_into(ts[0], ts[0], eigs, norb, nelec)
                if mats:
                    contract_diag_coulombs_in_place(ts[1:], mats, norb, nelec, z_rep)
                orbital_rotations_in_place(ts, [vecs] + rots, norb, nelec)
                for t in ts:
                    result.add_(t)
                return result
            t = torch.empty_like(v)
            t.copy_(v)
            orbital_rotation_in_place(t, vecs.T.conj(), norb, nelec)
            contract_num_op_sum_into(t, t, eigs, norb, nelec)
            orbital_rotation_in_place(t, vecs, norb, nelec)
            axpby_(result, result, t, 1.0, 1.0)
            for mat, rot in zip(mats, rots):
                t.copy_(v)
                orbital_rotation_in_place(t, rot.T.conj(), norb, nelec)
                contract_diag_coulomb_into(t, t, mat, norb, nelec, z_rep)
                orbital_rotation_in_place(t, rot, norb, nelec)
                axpby_(result, result, t, 1.0, 1.0)
            return result

        return DeviceLinearOperator(norb, nelec, matvec_device)
