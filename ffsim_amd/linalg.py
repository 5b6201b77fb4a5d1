"""Host-side linear algebra used by the drop-ins (python/ffsim/linalg/)."""

from __future__ import annotations

import numpy as np
import scipy.linalg

from ffsim_amd import _lib


def givens_decomposition(mat: np.ndarray, tol: float = 1e-12):
    r"""Exact Givens (Clements) decomposition :math:`U = D G_L^* \cdots G_1^*`.

    Same contract as ffsim.linalg.givens_decomposition without compression
    (python/ffsim/linalg/givens.py:59-156): a list of ``(c, s, i, j)`` with
    ``|i - j| == 1`` acting as ``[[c, s], [-conj(s), c]]`` on ``(i, j)``, in
    application order, and the diagonal of ``D``.  Computed in FP64 by the
    library's host routine.
    """
    mat = np.ascontiguousarray(np.asarray(mat), dtype=np.complex128)
    if mat.ndim != 2 or mat.shape[0] != mat.shape[1]:
        raise ValueError("mat must be a square matrix")
    n = mat.shape[0]
    cap = max(n * (n - 1) // 2, 1)
    c = np.zeros(cap)
    s = np.zeros(2 * cap)
    i = np.zeros(cap, dtype=np.int32)
    j = np.zeros(cap, dtype=np.int32)
    n_rot = np.zeros(1, dtype=np.int32)
    d = np.zeros(2 * max(n, 1))
    ptr = lambda a: a.ctypes.data_as(__import__("ctypes").c_void_p)  # noqa: E731
    _lib.call("ffsim_givens_decomposition", ptr(mat), n, float(tol), ptr(c), ptr(s), ptr(i), ptr(j),
              ptr(n_rot), ptr(d))
    k = int(n_rot[0])
    rots = [(float(c[t]), complex(s[2 * t], s[2 * t + 1]), int(i[t]), int(j[t])) for t in range(k)]
    return rots, d[: 2 * n].view(np.complex128).copy()


def is_hermitian(mat, *, rtol=1e-5, atol=1e-8) -> bool:
    m, n = mat.shape
    return m == n and np.allclose(mat, mat.T.conj(), rtol=rtol, atol=atol)


def is_real_symmetric(mat, *, rtol=1e-5, atol=1e-8) -> bool:
    m, n = mat.shape
    return m == n and bool(np.all(np.isreal(mat))) and np.allclose(mat, mat.T, rtol=rtol, atol=atol)


def is_unitary(mat, *, rtol=1e-5, atol=1e-8) -> bool:
    m, n = mat.shape
    return m == n and np.allclose(mat @ mat.T.conj(), np.eye(m), rtol=rtol, atol=atol)


def antihermitians_to_parameters(mats: np.ndarray) -> np.ndarray:
    """Real parameters of anti-Hermitian matrices (python/ffsim/linalg/util.py:67-90).

    Per matrix: Re of the strict upper triangle (row-major triu order), then
    Im of the upper triangle including the diagonal.
    """
    mats = np.asarray(mats)
    n_mats, dim, _ = mats.shape
    n_triu = dim * (dim - 1) // 2
    params = np.zeros((n_mats, dim * dim))
    r, c = np.triu_indices(dim, k=1)
    params[:, :n_triu] = mats[:, r, c].real
    r, c = np.triu_indices(dim)
    params[:, n_triu:] = mats[:, r, c].imag
    return params.reshape(-1)


def antihermitians_from_parameters(params: np.ndarray, dim: int, n_mats: int) -> np.ndarray:
    """Inverse of antihermitians_to_parameters (python/ffsim/linalg/util.py:93-124)."""
    params = np.asarray(params, dtype=float).reshape(n_mats, dim * dim)
    n_triu = dim * (dim - 1) // 2
    mats = np.zeros((n_mats, dim, dim), dtype=complex)
    r, c = np.triu_indices(dim)
    vals = 1j * params[:, n_triu:]
    mats[:, r, c] = vals
    mats[:, c, r] = vals
    r, c = np.triu_indices(dim, k=1)
    mats[:, r, c] += params[:, :n_triu]
    mats[:, c, r] -= params[:, :n_triu]
    return mats


def unitary_from_parameters(params: np.ndarray, dim: int) -> np.ndarray:
    """U = expm(K(params)) (python/ffsim/linalg/util.py:151-168)."""
    return scipy.linalg.expm(antihermitians_from_parameters(params, dim, 1)[0])


def unitary_to_parameters(mat: np.ndarray) -> np.ndarray:
    """Parameters of logm(U) (python/ffsim/linalg/util.py:127-148)."""
    return antihermitians_to_parameters(scipy.linalg.logm(np.asarray(mat))[None])
