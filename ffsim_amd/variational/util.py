"""Helpers shared by the UCJ constructors (python/ffsim/variational/util.py)."""

from __future__ import annotations

import numpy as np
import scipy.linalg


def orbital_rotation_from_t1_amplitudes(t1: np.ndarray) -> np.ndarray:
    r"""The orbital rotation :math:`e^{K}` with :math:`K_{ai} = t_{ia}`,
    :math:`K_{ia} = -t^*_{ia}` and zeros elsewhere (util.py:38-54)."""
    t1 = np.asarray(t1)
    nocc, nvrt = t1.shape
    norb = nocc + nvrt
    gen = np.zeros((norb, norb), dtype=t1.dtype)
    gen[nocc:, :nocc] = t1.T
    gen[:nocc, nocc:] = -t1.conj()
    return scipy.linalg.expm(gen)


def pair_mask(norb: int, pairs, *, symmetric: bool) -> np.ndarray:
    """Boolean norb x norb mask that is True on the listed ``(i, j)`` entries
    (and on ``(j, i)`` when ``symmetric``); an empty list masks everything."""
    mask = np.zeros((norb, norb), dtype=bool)
    for i, j in pairs:
        mask[i, j] = True
        if symmetric:
            mask[j, i] = True
    return mask


def pad_reps(mats: np.ndarray, rots: np.ndarray, n_reps: int | None):
    """Pad the repetitions up to ``n_reps`` with identities: zero diagonal
    Coulomb matrices and identity orbital rotations (ucj_spin_balanced.py:497-505,
    ucj_spin_unbalanced.py:772-794)."""
    n = mats.shape[0]
    if n_reps is None or n >= n_reps:
        return mats, rots
    norb = mats.shape[-1]
    mats = np.concatenate([mats, np.zeros((n_reps - n,) + mats.shape[1:])])
    eye = np.broadcast_to(np.eye(norb), (n_reps - n,) + rots.shape[1:])
    rots = np.concatenate([rots, eye])
    return mats, rots
