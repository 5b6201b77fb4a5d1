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
                "expected diag_coulomb_mats of shape (n_reps, 3, norb, norb); "
                f"got {self.diag_coulomb_mats.shape}"
            )
        if self.orbital_rotations.ndim != 4 or self.orbital_rotations.shape[1] != 2:
            raise ValueError(
                "expected orbital_rotations of shape (n_reps, 2, norb, norb); "
                f"got {self.orbital_rotations.shape}"
            )
        if self.final_orbital_rotation is not None and self.final_orbital_rotation.ndim != 3:
            raise ValueError(
                "expected final_orbital_rotation of shape (2, norb, norb); "
                f"got {self.final_orbital_rotation.shape}"
            )
        if self.diag_coulomb_mats.shape[0] != self.orbital_rotations.shape[0]:
            raise ValueError(
                f"diag_coulomb_mats and orbital_rotations hold {self.diag_coulomb_mats.shape[0]} and "
                f"{self.orbital_rotations.shape[0]} repetitions; they must match"
            )
        if not all(
            _linalg.is_real_symmetric(m[0], rtol=rtol, atol=atol)
            and _linalg.is_real_symmetric(m[2], rtol=rtol, atol=atol)
            for m in self.diag_coulomb_mats
        ):
            raise ValueError(
                "the same-spin (aa and bb) diagonal Coulomb matrices must be real symmetric"
            )
        if not all(
            _linalg.is_unitary(u[0], rtol=rtol, atol=atol)
            and _linalg.is_unitary(u[1], rtol=rtol, atol=atol)
            for u in self.orbital_rotations
        ):
            raise ValueError("every orbital rotation must be unitary")
        if self.final_orbital_rotation is not None and not (
            _linalg.is_unitary(self.final_orbital_rotation[0], rtol=rtol, atol=atol)
            and _linalg.is_unitary(self.final_orbital_rotation[1], rtol=rtol, atol=atol)
        ):
            raise ValueError("the final orbital rotation must be unitary")

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
                f"these settings take {n_params} parameters, got {len(params)}"
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

    @staticmethod
    def from_t_amplitudes(t2, *, t1=None, n_reps: int | tuple[int, int] | None = None,
                          interaction_pairs=None, tol: float = 1e-8, optimize: bool = False,
                          **optimize_kwargs) -> "UCJOpSpinUnbalanced":
        """Operator from ``t2 = (t2aa, t2ab, t2bb)`` (and ``t1 = (t1a, t1b)``)
        amplitudes (ucj_spin_unbalanced.py:430-703).

        The alpha-beta repetitions come first: the terms of the opposite-spin
        double factorization of ``t2ab``.  The same-spin ones follow, term ``k``
        pairing the ``k``-th terms of the factorizations of ``t2aa`` and
        ``t2bb`` (a missing one is the identity).  An int ``n_reps`` caps the
        total, alpha-beta terms first; a pair ``(n_ab, n_same)`` caps each kind
        separately; either pads with identities.  ``optimize=True`` raises
        ``NotImplementedError``.
        """
        from ffsim_amd.variational import util

        _linalg.check_optimize_kwargs(optimize_kwargs)
        if isinstance(n_reps, int) and n_reps <= 0:
            raise ValueError(f"n_reps must be positive, got {n_reps}")
        pairs_aa, pairs_ab, pairs_bb = (
            interaction_pairs if interaction_pairs is not None else (None, None, None)
        )
        _validate_interaction_pairs(pairs_aa, ordered=False)
        _validate_interaction_pairs(pairs_ab, ordered=True)
        _validate_interaction_pairs(pairs_bb, ordered=False)
        t2aa, t2ab, t2bb = t2
        nocc_a, _, nvrt_a, _ = np.shape(t2ab)
        norb = nocc_a + nvrt_a
        df = dict(tol=tol, optimize=optimize, **optimize_kwargs)

        def empty():
            return np.empty((0, 3, norb, norb)), np.empty((0, 2, norb, norb))

        if n_reps is None or isinstance(n_reps, int):
            cap_ab = n_reps
        else:
            cap_ab = n_reps[0]
        if cap_ab is None or cap_ab >= 1:
            mats_ab, rots_ab = _linalg.double_factorized_t2_alpha_beta(
                t2ab, max_terms=cap_ab, **df)
        elif cap_ab == 0:
            mats_ab, rots_ab = empty()
        else:
            raise ValueError(
                f"the alpha-beta (opposite-spin) part needs a positive number of repetitions, got {cap_ab}")

        if n_reps is None:
            cap_same = None
        elif isinstance(n_reps, int):
            cap_same = n_reps - len(mats_ab)
        else:
            cap_same = n_reps[1]
        if cap_same is None or cap_same >= 1:
            mats_aa, rots_aa = _linalg.double_factorized_t2(t2aa, max_terms=cap_same, **df)
            mats_bb, rots_bb = _linalg.double_factorized_t2(t2bb, max_terms=cap_same, **df)
            n_same = max(len(mats_aa), len(mats_bb))
            if n_same:
                mats_same = np.zeros((n_same, 3, norb, norb))
                mats_same[: len(mats_aa), 0] = mats_aa
                mats_same[: len(mats_bb), 2] = mats_bb
                eye = np.eye(norb)
                rots_same = np.stack([
                    [rots_aa[k] if k < len(rots_aa) else eye,
                     rots_bb[k] if k < len(rots_bb) else eye]
                    for k in range(n_same)])
            else:
                mats_same, rots_same = empty()
        elif cap_same == 0:
            mats_same, rots_same = empty()
        else:
            raise ValueError(
                f"the same-spin part needs a positive number of repetitions, got {cap_same}")

        if n_reps is None or isinstance(n_reps, int):
            mats, rots = util.pad_reps(np.concatenate([mats_ab, mats_same])[:n_reps],
                                       np.concatenate([rots_ab, rots_same])[:n_reps], n_reps)
        else:
            n_ab, n_same = n_reps
            mats_ab, rots_ab = util.pad_reps(mats_ab[:n_ab], rots_ab[:n_ab], n_ab)
            mats_same, rots_same = util.pad_reps(mats_same[:n_same], rots_same[:n_same],
                                                 n_same)
            mats = np.concatenate([mats_ab, mats_same])
            rots = np.concatenate([rots_ab, rots_same])

        final = None
        if t1 is not None:
            final = np.stack([util.orbital_rotation_from_t1_amplitudes(t) for t in t1])
        for kind, pairs, symmetric in ((0, pairs_aa, True), (1, pairs_ab, False),
                                       (2, pairs_bb, True)):
            if pairs is not None:
                mats[:, kind] *= util.pair_mask(norb, pairs, symmetric=symmetric)
        return UCJOpSpinUnbalanced(diag_coulomb_mats=mats, orbital_rotations=rots,
                                   final_orbital_rotation=final)

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
