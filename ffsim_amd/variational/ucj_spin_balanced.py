"""Spin-balanced (L)UCJ operator (python/ffsim/variational/ucj_spin_balanced.py).

The dataclass mirrors ffsim's: the same fields, validation, parameter
layout and ``_apply_unitary_`` semantics.  ``_apply_unitary_`` runs the whole
ansatz as one C-ABI call on the GPU (fused window passes for every orbital
rotation, one elementwise pass per diagonal Coulomb evolution).
"""

from __future__ import annotations

import ctypes

import itertools
from dataclasses import InitVar, dataclass

import numpy as np
import torch

from ffsim_amd import _device, _lib
from ffsim_amd import linalg as _linalg


def _validate_interaction_pairs(pairs, ordered: bool) -> None:
    """python/ffsim/variational/util.py validate_interaction_pairs."""
    if pairs is None:
        return
    if len(set(pairs)) != len(pairs):
        raise ValueError(f"interaction pairs contain a repeated pair: {pairs}.")
    if not ordered:
        for i, j in pairs:
            if i > j:
                raise ValueError(
                    f"same-spin (or spinless) interaction pair {(i, j)} is lower triangular; "
                    "give each pair as (i, j) with i <= j"
                )


@dataclass(frozen=True)
class UCJOpSpinBalanced:
    r"""A spin-balanced unitary cluster Jastrow operator
    :math:`\prod_k \mathcal{U}_k e^{i \mathcal{J}_k} \mathcal{U}_k^\dagger`."""

    diag_coulomb_mats: np.ndarray  # (n_reps, 2, norb, norb)
    orbital_rotations: np.ndarray  # (n_reps, norb, norb)
    final_orbital_rotation: np.ndarray | None = None
    validate: InitVar[bool] = True
    rtol: InitVar[float] = 1e-5
    atol: InitVar[float] = 1e-8

    def __post_init__(self, validate: bool, rtol: float, atol: float):
        if not validate:
            return
        if self.diag_coulomb_mats.ndim != 4 or self.diag_coulomb_mats.shape[1] != 2:
            raise ValueError(
                "expected diag_coulomb_mats of shape (n_reps, 2, norb, norb); "
                f"got {self.diag_coulomb_mats.shape}"
            )
        if self.orbital_rotations.ndim != 3:
            raise ValueError(
                "expected orbital_rotations of shape (n_reps, norb, norb); "
                f"got {self.orbital_rotations.shape}"
            )
        if self.final_orbital_rotation is not None and self.final_orbital_rotation.ndim != 2:
            raise ValueError(
                "expected final_orbital_rotation of shape (norb, norb); "
                f"got {self.final_orbital_rotation.shape}"
            )
        if self.diag_coulomb_mats.shape[0] != self.orbital_rotations.shape[0]:
            raise ValueError(
                f"diag_coulomb_mats and orbital_rotations hold {self.diag_coulomb_mats.shape[0]} and "
                f"{self.orbital_rotations.shape[0]} repetitions; they must match"
            )
        if not all(
            _linalg.is_real_symmetric(m[0], rtol=rtol, atol=atol)
            and _linalg.is_real_symmetric(m[1], rtol=rtol, atol=atol)
            for m in self.diag_coulomb_mats
        ):
            raise ValueError("every diagonal Coulomb matrix must be real symmetric")
        if not all(_linalg.is_unitary(u, rtol=rtol, atol=atol) for u in self.orbital_rotations):
            raise ValueError("every orbital rotation must be unitary")
        if self.final_orbital_rotation is not None and not _linalg.is_unitary(
            self.final_orbital_rotation, rtol=rtol, atol=atol
        ):
            raise ValueError("the final orbital rotation must be unitary")

    @property
    def norb(self) -> int:
        return self.diag_coulomb_mats.shape[-1]

    @property
    def n_reps(self) -> int:
        return self.diag_coulomb_mats.shape[0]

    @staticmethod
    def n_params(norb: int, n_reps: int, *, interaction_pairs=None,
                 with_final_orbital_rotation: bool = False) -> int:
        """Number of parameters (python/ffsim/variational/ucj_spin_balanced.py:143-195)."""
        pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
        _validate_interaction_pairs(pairs_aa, ordered=False)
        _validate_interaction_pairs(pairs_ab, ordered=False)
        n_triu = norb * (norb + 1) // 2
        n_aa = n_triu if pairs_aa is None else len(pairs_aa)
        n_ab = n_triu if pairs_ab is None else len(pairs_ab)
        return n_reps * (n_aa + n_ab + norb**2) + with_final_orbital_rotation * norb**2

    @staticmethod
    def from_parameters(params: np.ndarray, *, norb: int, n_reps: int, interaction_pairs=None,
                        with_final_orbital_rotation: bool = False) -> "UCJOpSpinBalanced":
        """Operator from a real parameter vector (ucj_spin_balanced.py:197-293)."""
        n_params = UCJOpSpinBalanced.n_params(
            norb, n_reps, interaction_pairs=interaction_pairs,
            with_final_orbital_rotation=with_final_orbital_rotation,
        )
        if len(params) != n_params:
            raise ValueError(
                f"these settings take {n_params} parameters, got {len(params)}"
            )
        pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
        triu = list(itertools.combinations_with_replacement(range(norb), 2))
        pairs_aa = triu if pairs_aa is None else pairs_aa
        pairs_ab = triu if pairs_ab is None else pairs_ab
        diag_coulomb_mats = np.zeros((n_reps, 2, norb, norb))
        orbital_rotations = np.zeros((n_reps, norb, norb), dtype=complex)
        index = 0
        for rep in range(n_reps):
            orbital_rotations[rep] = _linalg.unitary_from_parameters(
                params[index : index + norb**2], norb
            )
            index += norb**2
            for kind, pairs in enumerate((pairs_aa, pairs_ab)):
                if pairs:
                    rows, cols = zip(*pairs)
                    vals = params[index : index + len(pairs)]
                    diag_coulomb_mats[rep, kind][cols, rows] = vals
                    diag_coulomb_mats[rep, kind][rows, cols] = vals
                    index += len(pairs)
        final = None
        if with_final_orbital_rotation:
            final = _linalg.unitary_from_parameters(params[index:], norb)
        return UCJOpSpinBalanced(
            diag_coulomb_mats=diag_coulomb_mats,
            orbital_rotations=orbital_rotations,
            final_orbital_rotation=final,
        )

    @staticmethod
    def from_t_amplitudes(t2: np.ndarray, *, t1: np.ndarray | None = None,
                          n_reps: int | None = None, interaction_pairs=None,
                          tol: float = 1e-8, optimize: bool = False,
                          **optimize_kwargs) -> "UCJOpSpinBalanced":
        """Operator from t2 (and optionally t1) amplitudes
        (ucj_spin_balanced.py:377-540).

        The repetitions are the terms of the explicit double factorization of
        ``t2`` (same matrix for alpha-alpha and alpha-beta), cut to ``n_reps``
        or padded with identities up to it; ``t1`` gives the final orbital
        rotation; entries outside ``interaction_pairs`` are zeroed.
        ``optimize=True`` raises ``NotImplementedError``.
        """
        from ffsim_amd.variational import util

        _linalg.check_optimize_kwargs(optimize_kwargs)
        if isinstance(n_reps, int) and n_reps <= 0:
            raise ValueError(f"n_reps must be positive, got {n_reps}")
        pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
        _validate_interaction_pairs(pairs_aa, ordered=False)
        _validate_interaction_pairs(pairs_ab, ordered=False)
        mats, rots = _linalg.double_factorized_t2(t2, tol=tol, max_terms=n_reps,
                                                  optimize=optimize, **optimize_kwargs)
        norb = rots.shape[-1]
        mats, rots = util.pad_reps(np.stack([mats, mats], axis=1), rots, n_reps)
        for kind, pairs in enumerate((pairs_aa, pairs_ab)):
            if pairs is not None:
                mats[:, kind] *= util.pair_mask(norb, pairs, symmetric=True)
        final = None if t1 is None else util.orbital_rotation_from_t1_amplitudes(t1)
        return UCJOpSpinBalanced(diag_coulomb_mats=mats, orbital_rotations=rots,
                                 final_orbital_rotation=final)

    @staticmethod
    def from_cisd_vec(cisd_vec: np.ndarray, *, norb: int, nocc: int,
                      c0_threshold: float = 1e-12, n_reps: int | None = None,
                      interaction_pairs=None, tol: float = 1e-8, optimize: bool = False,
                      **optimize_kwargs) -> "UCJOpSpinBalanced":
        r"""Operator from a restricted CISD vector (ucj_spin_balanced.py:534-655).

        The vector holds :math:`c_0`, then the singles ``c1[i, a]`` and then the
        doubles ``c2[i, j, a, b]``, each flattened row-major (the layout of
        pyscf 2.14.0 ``ci.cisd.cisdvec_to_amplitudes`` for RCISD).  With
        :math:`t_1 = c_1 / c_0` and
        :math:`t_{2,ijab} = c_{2,ijab} / c_0 - t_{1,ia} t_{1,jb} / 2` the
        result is :meth:`from_t_amplitudes`.  Raises ``ValueError`` when
        :math:`|c_0|` is within ``c0_threshold`` of zero or the length does not
        fit ``norb`` and ``nocc``.
        """
        vec = np.asarray(cisd_vec)
        nvrt = norb - nocc
        n1 = nocc * nvrt
        if vec.ndim != 1 or vec.shape[0] != 1 + n1 + n1 * n1:
            raise ValueError(
                f"A CISD vector for norb={norb}, nocc={nocc} has {1 + n1 + n1 * n1} "
                f"entries. Got shape {vec.shape}.")
        c0 = vec[0]
        if abs(c0) <= c0_threshold:
            raise ValueError(
                f"CISD vector's reference amplitude c0={c0} is within the threshold "
                f"c0_tol={c0_threshold} of zero")
        t1 = vec[1 : 1 + n1].reshape(nocc, nvrt) / c0
        t2 = vec[1 + n1 :].reshape(nocc, nocc, nvrt, nvrt) / c0
        t2 = t2 - 0.5 * t1[:, None, :, None] * t1[None, :, None, :]
        return UCJOpSpinBalanced.from_t_amplitudes(
            t2, t1=t1, n_reps=n_reps, interaction_pairs=interaction_pairs, tol=tol,
            optimize=optimize, **optimize_kwargs)

    def to_parameters(self, *, interaction_pairs=None) -> np.ndarray:
        """Real parameter vector (ucj_spin_balanced.py:296-373)."""
        n_reps, _, norb, _ = self.diag_coulomb_mats.shape
        n_params = UCJOpSpinBalanced.n_params(
            norb, n_reps, interaction_pairs=interaction_pairs,
            with_final_orbital_rotation=self.final_orbital_rotation is not None,
        )
        pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
        triu = list(itertools.combinations_with_replacement(range(norb), 2))
        pairs_aa = triu if pairs_aa is None else pairs_aa
        pairs_ab = triu if pairs_ab is None else pairs_ab
        params = np.zeros(n_params)
        index = 0
        for rot, mats in zip(self.orbital_rotations, self.diag_coulomb_mats):
            params[index : index + norb**2] = _linalg.unitary_to_parameters(rot)
            index += norb**2
            for pairs, m in zip((pairs_aa, pairs_ab), mats):
                if pairs:
                    params[index : index + len(pairs)] = m[tuple(zip(*pairs))]
                    index += len(pairs)
        if self.final_orbital_rotation is not None:
            params[index:] = _linalg.unitary_to_parameters(self.final_orbital_rotation)
        return params

    def apply_in_place(self, t: torch.Tensor, norb: int, nelec) -> torch.Tensor:
        """Apply to a resident 1-D device state in place."""
        na, nb = nelec
        _device.check_size(t, norb, nelec)
        dcm = np.ascontiguousarray(self.diag_coulomb_mats, dtype=float)
        rots = np.ascontiguousarray(self.orbital_rotations, dtype=np.complex128)
        if dcm.shape[-1] != norb or rots.shape[-1] != norb:
            raise ValueError(f"operator has norb={self.norb}, expected {norb}")
        final = (
            np.ascontiguousarray(self.final_orbital_rotation, dtype=np.complex128)
            if self.final_orbital_rotation is not None
            else None
        )
        work = _device.work_buffer(t, norb, na, nb)
        _lib.call(
            "ffsim_apply_ucj_spin_balanced",
            _device.ptr(t),
            _device.ptr(work),
            _device.dtype_tag(t),
            norb,
            na,
            nb,
            self.n_reps,
            _device.host_ptr(dcm),
            _device.host_ptr(rots),
            _device.host_ptr(final),
            _device.stream_ptr(),
        )
        return t

    @staticmethod
    def apply_batched_in_place(ops, states, norb: int, nelec, works=None) -> list:
        """Apply ``ops[i]`` to the resident 1-D device state ``states[i]`` in place, for
        states of one shape and dtype, in one C-ABI call (batched launches for small
        states).  ``works``: optional per-state device buffers of the state's size
        and dtype; allocated here when not given.  Operators must share ``n_reps``
        and whether they have a final orbital rotation."""
        ops, states = list(ops), list(states)
        if len(ops) != len(states):
            raise ValueError("one operator per state")
        if not states:
            return states
        na, nb = nelec
        n_reps = ops[0].n_reps
        has_final = ops[0].final_orbital_rotation is not None
        tag = _device.dtype_tag(states[0])
        for op, t in zip(ops, states):
            _device.check_size(t, norb, nelec)
            if op.n_reps != n_reps or (op.final_orbital_rotation is not None) != has_final:
                raise ValueError("a batch needs operators with the same n_reps and final-rotation form")
            if op.norb != norb:
                raise ValueError(f"operator has norb={op.norb}, expected {norb}")
            if _device.dtype_tag(t) != tag or t.device != states[0].device:
                raise ValueError("a batch needs states of one dtype on one device")
        if works is None:
            works = [torch.empty_like(t) for t in states]
        dcm = np.ascontiguousarray(np.stack([op.diag_coulomb_mats for op in ops]), dtype=float)
        rots = np.ascontiguousarray(np.stack([op.orbital_rotations for op in ops]),
                                    dtype=np.complex128)
        final = (np.ascontiguousarray(np.stack([op.final_orbital_rotation for op in ops]),
                                      dtype=np.complex128) if has_final else None)
        vec_ptrs = (ctypes.c_void_p * len(states))(*[t.data_ptr() for t in states])
        work_ptrs = (ctypes.c_void_p * len(states))(
            *[w.data_ptr() if w is not None else None for w in works])
        _lib.call("ffsim_apply_ucj_spin_balanced_batched", len(states), vec_ptrs, work_ptrs,
                  tag, norb, na, nb, n_reps, _device.host_ptr(dcm), _device.host_ptr(rots),
                  _device.host_ptr(final), _device.stream_ptr())
        return states

    def _apply_unitary_(self, vec, norb: int, nelec, copy: bool):
        if isinstance(nelec, int):
            return NotImplemented
        st = _device.to_device(vec, copy)
        self.apply_in_place(st.tensor, norb, nelec)
        return st.restore(st.tensor)

    def _approx_eq_(self, other, rtol: float, atol: float) -> bool:
        if isinstance(other, UCJOpSpinBalanced):
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
