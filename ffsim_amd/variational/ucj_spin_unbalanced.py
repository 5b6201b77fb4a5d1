"""Spin-unbalanced (L)UCJ operator (python/ffsim/variational/ucj_spin_unbalanced.py).

Separate alpha and beta orbital rotations and three diagonal Coulomb
matrices (aa, ab, bb) per repetition.  ``_apply_unitary_`` follows
ucj_spin_unbalanced.py:705-745: per repetition one two-sector orbital
rotation (U_k^dag U_{k-1} per spin) and one diagonal Coulomb evolution at
time -1, then the final rotation, each one C-ABI call on the resident state.
"""

from __future__ import annotations

import itertools
from dataclasses import InitVar, dataclass

import numpy as np

from ffsim_amd import _device
from ffsim_amd import linalg as _linalg
from ffsim_amd.gates.diag_coulomb import diag_coulomb_evolution_in_place
from ffsim_amd.gates.orbital_rotation import orbital_rotation_in_place
from ffsim_amd.variational.ucj_spin_balanced import _validate_interaction_pairs


@dataclass(frozen=True)
class UCJOpSpinUnbalanced:
    r"""A spin-unbalanced unitary cluster Jastrow operator."""

    diag_coulomb_mats: np.ndarray  # (n_reps, 3, norb, norb)
    orbital_rotations: np.ndarray  # (n_reps, 2, norb, norb)
    final_orbital_rotation: np.ndarray | None = None  # (2, norb, norb)
    validate: InitVar[bool] = True
    rtol: InitVar[float] = 1e-5
    atol: InitVar[float] = 1e-8

    def __post_init__(self, validate: bool, rtol: float, atol: float):
        """Validation of ucj_spin_unbalanced.py:85-133."""
        if not validate:
            return
        if self.diag_coulomb_mats.ndim != 4 or self.diag_coulomb_mats.shape[1] != 3:
            raise ValueError(
                "diag_coulomb_mats should have shape (n_reps, 3, norb, norb). "
                f"Got shape {self.diag_coulomb_mats.shape}."
            )
        if self.orbital_rotations.ndim != 4 or self.orbital_rotations.shape[1] != 2:
            raise ValueError(
                "orbital_rotations should have shape (n_reps, 2, norb, norb). "
                f"Got shape {self.orbital_rotations.shape}."
            )
        if self.final_orbital_rotation is not None and self.final_orbital_rotation.ndim != 3:
            raise ValueError(
                "final_orbital_rotation should have shape (2, norb, norb). "
                f"Got shape {self.final_orbital_rotation.shape}."
            )
        if self.diag_coulomb_mats.shape[0] != self.orbital_rotations.shape[0]:
            raise ValueError(
                "diag_coulomb_mats and orbital_rotations should have the same first "
                f"dimension. Got {self.diag_coulomb_mats.shape[0]} and "
                f"{self.orbital_rotations.shape[0]}."
            )
        if not all(
            _linalg.is_real_symmetric(m[0], rtol=rtol, atol=atol)
            and _linalg.is_real_symmetric(m[2], rtol=rtol, atol=atol)
            for m in self.diag_coulomb_mats
        ):
            raise ValueError(
                "alpha-alpha and beta-beta diagonal Coulomb matrices were not all "
                "real symmetric."
            )
        if not all(
            _linalg.is_unitary(u[0], rtol=rtol, atol=atol)
            and _linalg.is_unitary(u[1], rtol=rtol, atol=atol)
            for u in self.orbital_rotations
        ):
            raise ValueError("Orbital rotations were not all unitary.")
        if self.final_orbital_rotation is not None and not (
            _linalg.is_unitary(self.final_orbital_rotation[0], rtol=rtol, atol=atol)
            and _linalg.is_unitary(self.final_orbital_rotation[1], rtol=rtol, atol=atol)
        ):
            raise ValueError("Final orbital rotation was not unitary.")

    @property
    def norb(self) -> int:
        return self.diag_coulomb_mats.shape[-1]

    @property
    def n_reps(self) -> int:
        return self.diag_coulomb_mats.shape[0]

    @staticmethod
    def _pairs(norb, interaction_pairs):
        pairs_aa, pairs_ab, pairs_bb = (
            interaction_pairs if interaction_pairs is not None else (None, None, None)
        )
        _validate_interaction_pairs(pairs_aa, ordered=False)
        _validate_interaction_pairs(pairs_ab, ordered=True)
        _validate_interaction_pairs(pairs_bb, ordered=False)
        triu = list(itertools.combinations_with_replacement(range(norb), 2))
        full = list(itertools.product(range(norb), repeat=2))
        return (
            triu if pairs_aa is None else list(pairs_aa),
            full if pairs_ab is None else list(pairs_ab),
            triu if pairs_bb is None else list(pairs_bb),
        )

    @staticmethod
    def n_params(norb: int, n_reps: int, *, interaction_pairs=None,
                 with_final_orbital_rotation: bool = False) -> int:
        """ucj_spin_unbalanced.py:146-206."""
        pairs_aa, pairs_ab, pairs_bb = UCJOpSpinUnbalanced._pairs(norb, interaction_pairs)
        return (
            n_reps * (len(pairs_aa) + len(pairs_ab) + len(pairs_bb) + 2 * norb**2)
            + with_final_orbital_rotation * 2 * norb**2
        )

    @staticmethod
    def from_parameters(params: np.ndarray, *, norb: int, n_reps: int, interaction_pairs=None,
                        with_final_orbital_rotation: bool = False) -> "UCJOpSpinUnbalanced":
        """ucj_spin_unbalanced.py:208-329: per repetition the alpha then beta
        rotation generators, then the aa, ab, bb pair values."""
        n_params = UCJOpSpinUnbalanced.n_params(
            norb, n_reps, interaction_pairs=interaction_pairs,
            with_final_orbital_rotation=with_final_orbital_rotation,
        )
        if len(params) != n_params:
            raise ValueError(
                "The number of parameters passed did not match the number expected "
                "based on the function inputs. "
                f"Expected {n_params} but got {len(params)}."
            )
        pairs_aa, pairs_ab, pairs_bb = UCJOpSpinUnbalanced._pairs(norb, interaction_pairs)
        mats = np.zeros((n_reps, 3, norb, norb))
        rots = np.zeros((n_reps, 2, norb, norb), dtype=complex)
        index = 0
        for rep in range(n_reps):
            for sigma in range(2):
                rots[rep, sigma] = _linalg.unitary_from_parameters(
                    params[index : index + norb**2], norb
                )
                index += norb**2
            for kind, pairs, symmetric in ((0, pairs_aa, True), (1, pairs_ab, False),
                                           (2, pairs_bb, True)):
                if pairs:
                    rows, cols = zip(*pairs)
                    vals = params[index : index + len(pairs)]
                    if symmetric:
                        mats[rep, kind][cols, rows] = vals
                    mats[rep, kind][rows, cols] = vals
                    index += len(pairs)
        final = None
        if with_final_orbital_rotation:
            final = np.zeros((2, norb, norb), dtype=complex)
            for sigma in range(2):
                final[sigma] = _linalg.unitary_from_parameters(
                    params[index : index + norb**2], norb
                )
                index += norb**2
        return UCJOpSpinUnbalanced(
            diag_coulomb_mats=mats, orbital_rotations=rots, final_orbital_rotation=final
        )

    def to_parameters(self, interaction_pairs=None) -> np.ndarray:
        """ucj_spin_unbalanced.py:331-428."""
        n_reps, _, norb, _ = self.diag_coulomb_mats.shape
        n_params = UCJOpSpinUnbalanced.n_params(
            norb, n_reps, interaction_pairs=interaction_pairs,
            with_final_orbital_rotation=self.final_orbital_rotation is not None,
        )
        pairs = UCJOpSpinUnbalanced._pairs(norb, interaction_pairs)
        params = np.zeros(n_params)
        index = 0
        for rots, mats in zip(self.orbital_rotations, self.diag_coulomb_mats):
            for u in rots:
                params[index : index + norb**2] = _linalg.unitary_to_parameters(u)
                index += norb**2
            for p, m in zip(pairs, mats):
                if p:
                    params[index : index + len(p)] = m[tuple(zip(*p))]
                    index += len(p)
        if self.final_orbital_rotation is not None:
            for u in self.final_orbital_rotation:
                params[index : index + norb**2] = _linalg.unitary_to_parameters(u)
                index += norb**2
        return params

    def apply_in_place(self, t, norb: int, nelec):
        """Apply to a resident 1-D device state in place."""
        eye = np.eye(norb, dtype=complex)
        current = (eye, eye)
        for mats, rots in zip(self.diag_coulomb_mats, self.orbital_rotations):
            step = (rots[0].conj().T @ current[0], rots[1].conj().T @ current[1])
            orbital_rotation_in_place(t, step, norb, nelec)
            diag_coulomb_evolution_in_place(t, (mats[0], mats[1], mats[2]), -1.0, norb, nelec,
                                            False)
            current = (rots[0], rots[1])
        if self.final_orbital_rotation is not None:
            current = (self.final_orbital_rotation[0] @ current[0],
                       self.final_orbital_rotation[1] @ current[1])
        orbital_rotation_in_place(t, current, norb, nelec)
        return t

    def _apply_unitary_(self, vec, norb: int, nelec, copy: bool):
        if isinstance(nelec, int):
            return NotImplemented
        st = _device.to_device(vec, copy)
        self.apply_in_place(st.tensor, norb, tuple(nelec))
        return st.restore(st.tensor)

    def _approx_eq_(self, other, rtol: float, atol: float) -> bool:
        if isinstance(other, UCJOpSpinUnbalanced):
            if not np.allclose(self.diag_coulomb_mats, other.diag_coulomb_mats, rtol=rtol, atol=atol):
                return False
            if not np.allclose(self.orbital_rotations, other.orbital_rotations, rtol=rtol, atol=atol):
                return False
            if (self.final_orbital_rotation is None) != (other.final_orbital_rotation is None):
                return False
            if self.final_orbital_rotation is not None:
                return np.allclose(self.final_orbital_rotation, other.final_orbital_rotation,
                                   rtol=rtol, atol=atol)
            return True
        return NotImplemented
