"""Spin-balanced UCJ operator parameterized by gate angles
(python/ffsim/variational/ucj_angles_spin_balanced.py).

Each repetition is a Givens ansatz (an orbital rotation) and a
number-number ansatz (a diagonal Coulomb evolution).  ``_apply_unitary_``
(ucj_angles_spin_balanced.py:183-219) merges consecutive rotations as
U_k^dag U_{k-1} and runs on the resident state through the GPU drop-ins.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from ffsim_amd import _device
from ffsim_amd.gates.diag_coulomb import diag_coulomb_evolution_in_place
from ffsim_amd.gates.orbital_rotation import orbital_rotation_in_place
from ffsim_amd.variational.givens import GivensAnsatzOp, brickwork
from ffsim_amd.variational.num_num import NumNumAnsatzOpSpinBalanced


@dataclass(frozen=True)
class UCJAnglesOpSpinBalanced:
    """A spin-balanced UCJ operator parameterized by gate rotation angles."""

    norb: int
    num_num_ansatz_ops: list
    givens_ansatz_ops: list
    final_givens_ansatz_op: GivensAnsatzOp | None = None

    def __post_init__(self):
        if len(self.num_num_ansatz_ops) != len(self.givens_ansatz_ops):
            raise ValueError(
                f"{len(self.num_num_ansatz_ops)} number-number layers and "
                f"{len(self.givens_ansatz_ops)} Givens layers; each repetition needs one of each"
            )

    @property
    def n_reps(self) -> int:
        return len(self.num_num_ansatz_ops)

    @staticmethod
    def n_params(norb: int, n_reps: int, num_num_interaction_pairs, givens_interaction_pairs,
                 with_final_givens_ansatz_op: bool = False) -> int:
        """ucj_angles_spin_balanced.py:50-65."""
        pairs_aa, pairs_ab = num_num_interaction_pairs
        per_rep = len(pairs_aa) + len(pairs_ab) + 2 * len(givens_interaction_pairs) + norb
        return n_reps * per_rep + with_final_givens_ansatz_op * norb**2

    @staticmethod
    def from_parameters(params: np.ndarray, *, norb: int, n_reps: int, num_num_interaction_pairs,
                        givens_interaction_pairs,
                        with_final_givens_ansatz_op: bool = False) -> "UCJAnglesOpSpinBalanced":
        """Per repetition: Givens angles (thetas, phis, phases), then number-number
        angles; optionally a final full-brickwork Givens ansatz
        (ucj_angles_spin_balanced.py:67-128)."""
        n_params = UCJAnglesOpSpinBalanced.n_params(
            norb, n_reps, num_num_interaction_pairs, givens_interaction_pairs,
            with_final_givens_ansatz_op,
        )
        if len(params) != n_params:
            raise ValueError(
                f"these settings take {n_params} parameters, got {len(params)}"
            )
        n_nn = sum(len(p) for p in num_num_interaction_pairs)
        n_gv = 2 * len(givens_interaction_pairs) + norb
        num_num, givens = [], []
        index = 0
        for _ in range(n_reps):
            givens.append(GivensAnsatzOp.from_parameters(
                params[index : index + n_gv], norb=norb, interaction_pairs=givens_interaction_pairs))
            index += n_gv
            num_num.append(NumNumAnsatzOpSpinBalanced.from_parameters(
                params[index : index + n_nn], norb=norb,
                interaction_pairs=num_num_interaction_pairs))
            index += n_nn
        final = None
        if with_final_givens_ansatz_op:
            final = GivensAnsatzOp.from_parameters(
                params[index:], norb=norb, interaction_pairs=list(brickwork(norb, norb)))
        return UCJAnglesOpSpinBalanced(norb, num_num_ansatz_ops=num_num, givens_ansatz_ops=givens,
                                       final_givens_ansatz_op=final)

    def to_parameters(self) -> np.ndarray:
        parts = []
        for gv, nn in zip(self.givens_ansatz_ops, self.num_num_ansatz_ops):
            parts.append(gv.to_parameters())
            parts.append(nn.to_parameters())
        if self.final_givens_ansatz_op is not None:
            parts.append(self.final_givens_ansatz_op.to_parameters())
        return np.concatenate(parts) if parts else np.zeros(0)

    @staticmethod
    def from_t_amplitudes(t2: np.ndarray, *, t1: np.ndarray | None = None,
                          n_reps: int | None = None, interaction_pairs=None,
                          tol: float = 1e-8) -> "UCJAnglesOpSpinBalanced":
        """The spin-balanced UCJ operator from t2 (and t1) amplitudes, in angle form
        (ucj_angles_spin_balanced.py:143-158)."""
        from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

        return UCJAnglesOpSpinBalanced.from_ucj_op(UCJOpSpinBalanced.from_t_amplitudes(
            t2, t1=t1, n_reps=n_reps, interaction_pairs=interaction_pairs, tol=tol))

    @staticmethod
    def from_ucj_op(ucj_op) -> "UCJAnglesOpSpinBalanced":
        """From a matrix-based UCJOpSpinBalanced (ucj_angles_spin_balanced.py:161-181)."""
        num_num = [NumNumAnsatzOpSpinBalanced.from_diag_coulomb_mats(m)
                   for m in ucj_op.diag_coulomb_mats]
        givens = [GivensAnsatzOp.from_orbital_rotation(u) for u in ucj_op.orbital_rotations]
        final = None
        if ucj_op.final_orbital_rotation is not None:
            final = GivensAnsatzOp.from_orbital_rotation(ucj_op.final_orbital_rotation)
        return UCJAnglesOpSpinBalanced(norb=ucj_op.norb, num_num_ansatz_ops=num_num,
                                       givens_ansatz_ops=givens, final_givens_ansatz_op=final)

    def apply_in_place(self, t, norb: int, nelec):
        current = np.eye(norb, dtype=complex)
        for nn, gv in zip(self.num_num_ansatz_ops, self.givens_ansatz_ops):
            u = gv.to_orbital_rotation()
            orbital_rotation_in_place(t, u.conj().T @ current, norb, nelec)
            mat_aa, mat_ab = nn.to_diag_coulomb_mats()
            diag_coulomb_evolution_in_place(t, (mat_aa, mat_ab, mat_aa), -1.0, norb, nelec, False)
            current = u
        if self.final_givens_ansatz_op is not None:
            current = self.final_givens_ansatz_op.to_orbital_rotation() @ current
        orbital_rotation_in_place(t, current, norb, nelec)
        return t

    def _apply_unitary_(self, vec, norb: int, nelec, copy: bool):
        if isinstance(nelec, int):
            return NotImplemented
        st = _device.to_device(vec, copy)
        self.apply_in_place(st.tensor, norb, tuple(nelec))
        return st.restore(st.tensor)
