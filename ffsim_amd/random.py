"""Seeded input generators with ffsim's draw order (python/ffsim/random/random.py).

The same seed gives the same objects as ffsim's generators, so workloads
can be described by seed alone.
"""

from __future__ import annotations

import math

import numpy as np

from ffsim_amd.hamiltonians import DiagonalCoulombHamiltonian, DoubleFactorizedHamiltonian
from ffsim_amd.variational import UCJOpSpinBalanced


def random_state_vector(dim: int, *, seed=None, dtype=complex) -> np.ndarray:
    """random.py:24-47: N(0,1) real parts, then imaginary parts, normalized."""
    if dim < 1:
        raise ValueError("the dimension must be positive")
    rng = np.random.default_rng(seed)
    vec = rng.standard_normal(dim).astype(dtype, copy=False)
    if np.issubdtype(dtype, np.complexfloating):
        vec += 1j * rng.standard_normal(dim).astype(dtype, copy=False)
    vec /= np.linalg.norm(vec)
    return vec


def random_unitary(dim: int, *, seed=None, dtype=complex) -> np.ndarray:
    """Haar unitary via QR with phase fix (random.py:84-106)."""
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((dim, dim)).astype(dtype, copy=False)
    z += 1j * rng.standard_normal((dim, dim)).astype(dtype, copy=False)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def random_orthogonal(dim: int, seed=None, dtype=float) -> np.ndarray:
    """Haar orthogonal (random.py:109-130)."""
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((dim, dim)).astype(dtype, copy=False)
    q, r = np.linalg.qr(m)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def random_hermitian(dim: int, *, seed=None, dtype=complex) -> np.ndarray:
    """random.py:151-167."""
    rng = np.random.default_rng(seed)
    mat = rng.standard_normal((dim, dim)).astype(dtype, copy=False)
    mat += 1j * rng.standard_normal((dim, dim)).astype(dtype, copy=False)
    return mat + mat.T.conj()


def random_real_symmetric_matrix(dim: int, *, rank: int | None = None, seed=None, dtype=float):
    """random.py:170-189."""
    rng = np.random.default_rng(seed)
    if rank is None:
        rank = dim
    mat = rng.standard_normal((dim, rank)).astype(dtype, copy=False)
    return mat @ mat.T


def random_t2_amplitudes(norb: int, nocc: int, *, seed=None, dtype=complex) -> np.ndarray:
    """random.py:240-276: one N(0,1) draw per unordered pair of excitations
    ``(i, a) <= (j, b)`` (pairs ranked occupied-major), stored with the
    symmetry ``t2[i, j, a, b] == t2[j, i, b, a]``; for a complex dtype the
    tensor is then rotated by a Haar unitary drawn next from the same stream."""
    rng = np.random.default_rng(seed)
    nvrt = norb - nocc
    t2 = np.zeros((nocc, nocc, nvrt, nvrt), dtype=dtype)
    exc = [(i, a) for i in range(nocc) for a in range(nvrt)]
    for p in range(len(exc)):
        i, a = exc[p]
        for j, b in exc[p:]:
            t2[i, j, a, b] = t2[j, i, b, a] = rng.standard_normal()
    if np.issubdtype(dtype, np.complexfloating):
        full = np.zeros((norb,) * 4, dtype=dtype)
        full[:nocc, :nocc, nocc:, nocc:] = t2
        u = random_unitary(norb, seed=rng)
        full = np.einsum("ijab,iI,jJ,aA,bB->IJAB", full, u.conj(), u.conj(), u, u)
        t2 = full[:nocc, :nocc, nocc:, nocc:]
    return t2


def _symmetric_uniform(dim: int, *, mean: float, scale: float, seed=None) -> np.ndarray:
    """random.py:535-546."""
    rng = np.random.default_rng(seed)
    vals = mean + rng.uniform(-0.5 * scale, 0.5 * scale, size=dim * (dim + 1) // 2)
    mat = np.zeros((dim, dim))
    rows, cols = np.triu_indices(dim)
    mat[rows, cols] = vals
    mat[cols, rows] = vals
    return mat


def _symmetric_normal(dim: int, *, mean: float, scale: float, seed=None) -> np.ndarray:
    """random.py:549-560."""
    rng = np.random.default_rng(seed)
    vals = rng.normal(loc=mean, scale=scale, size=dim * (dim + 1) // 2)
    mat = np.zeros((dim, dim))
    rows, cols = np.triu_indices(dim)
    mat[rows, cols] = vals
    mat[cols, rows] = vals
    return mat


def random_ucj_op_spin_balanced(
    norb: int,
    *,
    n_reps: int = 1,
    interaction_pairs=None,
    with_final_orbital_rotation: bool = False,
    diag_coulomb_mean: float = 0.0,
    diag_coulomb_scale: float = 2 * math.pi,
    diag_coulomb_normal: bool = False,
    seed=None,
) -> UCJOpSpinBalanced:
    """random.py:563-665."""
    pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
    rng = np.random.default_rng(seed)
    gen = _symmetric_normal if diag_coulomb_normal else _symmetric_uniform
    diag_coulomb_mats = np.stack(
        [
            np.stack(
                [
                    gen(norb, mean=diag_coulomb_mean, scale=diag_coulomb_scale, seed=rng),
                    gen(norb, mean=diag_coulomb_mean, scale=diag_coulomb_scale, seed=rng),
                ]
            )
            for _ in range(n_reps)
        ]
    ) if n_reps else np.zeros((0, 2, norb, norb))
    orbital_rotations = (
        np.stack([random_unitary(norb, seed=rng) for _ in range(n_reps)])
        if n_reps
        else np.zeros((0, norb, norb), dtype=complex)
    )
    final = random_unitary(norb, seed=rng) if with_final_orbital_rotation else None
    for kind, pairs in ((0, pairs_aa), (1, pairs_ab)):
        if pairs is not None:
            mask = np.zeros((norb, norb), dtype=bool)
            if pairs:
                rows, cols = zip(*pairs)
                mask[rows, cols] = True
                mask[cols, rows] = True
            diag_coulomb_mats[:, kind] *= mask
    return UCJOpSpinBalanced(
        diag_coulomb_mats=diag_coulomb_mats,
        orbital_rotations=orbital_rotations,
        final_orbital_rotation=final,
    )


def random_diagonal_coulomb_hamiltonian(norb: int, *, real: bool = False, seed=None):
    """random.py:883-908."""
    rng = np.random.default_rng(seed)
    one_body = (
        random_real_symmetric_matrix(norb, seed=rng) if real else random_hermitian(norb, seed=rng)
    )
    mat_aa = random_real_symmetric_matrix(norb, seed=rng)
    mat_ab = random_real_symmetric_matrix(norb, seed=rng)
    constant = rng.standard_normal()
    return DiagonalCoulombHamiltonian(one_body, np.stack([mat_aa, mat_ab]), constant=constant)


def random_double_factorized_hamiltonian(
    norb: int, *, rank: int | None = None, z_representation: bool = False, real: bool = False,
    seed=None,
) -> DoubleFactorizedHamiltonian:
    """random.py:911-956."""
    if rank is None:
        rank = norb * (norb + 1) // 2
    rng = np.random.default_rng(seed)
    if real:
        one_body = random_real_symmetric_matrix(norb, seed=rng)
        rots = np.stack([random_orthogonal(norb, seed=rng) for _ in range(rank)])
    else:
        one_body = random_hermitian(norb, seed=rng)
        rots = np.stack([random_unitary(norb, seed=rng) for _ in range(rank)])
    mats = np.stack([random_real_symmetric_matrix(norb, seed=rng) for _ in range(rank)])
    constant = rng.standard_normal()
    return DoubleFactorizedHamiltonian(
        one_body_tensor=one_body,
        diag_coulomb_mats=mats,
        orbital_rotations=rots,
        constant=constant,
        z_representation=z_representation,
    )


def random_df_hamiltonian_normal(
    norb: int, *, n_factors: int = 8, scale: float = 0.1, seed=None
) -> DoubleFactorizedHamiltonian:
    """Synthetic double-factorized Hamiltonian for energy-gradient workloads.

    One-body tensor: real symmetric with N(0, scale) entries; then n_factors
    Haar orbital rotations; then n_factors real symmetric N(0, scale)
    diagonal Coulomb matrices (number representation, constant 0).  The
    rotation-then-matrix draw order follows random.py:911-956; ffsim has no
    generator with these distributions.
    """
    rng = np.random.default_rng(seed)
    one_body = _symmetric_normal(norb, mean=0.0, scale=scale, seed=rng)
    rots = np.stack([random_unitary(norb, seed=rng) for _ in range(n_factors)])
    mats = np.stack([_symmetric_normal(norb, mean=0.0, scale=scale, seed=rng)
                     for _ in range(n_factors)])
    return DoubleFactorizedHamiltonian(
        one_body_tensor=one_body, diag_coulomb_mats=mats, orbital_rotations=rots, constant=0.0
    )
