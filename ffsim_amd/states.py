"""State dimensions and reference states (python/ffsim/states/)."""

from __future__ import annotations

import math

import numpy as np


def dims(norb: int, nelec: tuple[int, int]) -> tuple[int, int]:
    """(dim_a, dim_b) (python/ffsim/states/dimensions.py:18-33)."""
    n_alpha, n_beta = nelec
    return math.comb(norb, n_alpha), math.comb(norb, n_beta)


def dim(norb: int, nelec) -> int:
    """Dimension of the FCI space (python/ffsim/states/dimensions.py:36-50)."""
    if isinstance(nelec, int):
        return math.comb(norb, nelec)
    n_alpha, n_beta = nelec
    return math.comb(norb, n_alpha) * math.comb(norb, n_beta)


def hartree_fock_state(norb: int, nelec) -> np.ndarray:
    """One-hot state at address 0, i.e. the lowest orbitals occupied
    (python/ffsim/states/slater.py:122-138)."""
    vec = np.zeros(dim(norb, nelec), dtype=complex)
    vec[0] = 1
    return vec
