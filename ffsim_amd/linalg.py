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


# ---------------------------------------------------------------------------
# Double factorization of t2 amplitudes (the explicit, nested-eigendecomposition
# form).  The optimized ("compressed") form needs an autodiff framework that is
# not part of this package and raises NotImplementedError.


def _ranked_eigh(mat: np.ndarray, tol: float):
    """Eigenpairs of a real symmetric matrix, largest |eigenvalue| first, with the
    tail whose summed |eigenvalue| stays below ``tol`` dropped
    (python/ffsim/linalg/double_factorized_decomposition.py:39-53)."""
    vals, vecs = scipy.linalg.eigh(mat)
    order = np.argsort(np.abs(vals))[::-1]
    vals, vecs = vals[order], vecs[:, order]
    n_drop = int(np.searchsorted(np.cumsum(np.abs(vals[::-1])), tol))
    return vals[: len(vals) - n_drop], vecs[:, : len(vals) - n_drop]


def _ranked_svd(mat: np.ndarray, tol: float):
    """Thin SVD with the tail whose summed singular values stay below ``tol``
    dropped (double_factorized_decomposition.py:56-67)."""
    left, sing, right = scipy.linalg.svd(mat, full_matrices=False)
    keep = len(sing) - int(np.searchsorted(np.cumsum(sing[::-1]), tol))
    return left[:, :keep], sing[:keep], right[:keep]


def _excitation_matrix(vec: np.ndarray, nocc: int, norb: int) -> np.ndarray:
    """norb x norb real matrix with ``vec[i * nvrt + a]`` at row ``nocc + a``,
    column ``i`` and zeros elsewhere (the occupied->virtual block)."""
    mat = np.zeros((norb, norb))
    mat[nocc:, :nocc] = np.asarray(vec).reshape(nocc, norb - nocc).T
    return mat


def _quadrature(mat: np.ndarray, sign: int) -> np.ndarray:
    """Hermitian matrix (1 - s i)/2 (M + s i M^dagger)
    (double_factorized_decomposition.py:1216-1217)."""
    return 0.5 * (1 - sign * 1j) * (mat + sign * 1j * mat.T.conj())


# Keyword arguments of ffsim's t2 factorizations that only the optimized form
# reads (double_factorized_decomposition.py:467-482).
OPTIMIZE_ARGS = frozenset({"method", "callback", "options", "regularization",
                           "multi_stage_start", "multi_stage_step"})
_T2_ARGS = OPTIMIZE_ARGS | {"diag_coulomb_indices", "return_optimize_result"}


def check_optimize_kwargs(kwargs: dict, allowed=OPTIMIZE_ARGS) -> None:
    """TypeError for a keyword argument outside ``allowed``, as a signature that
    spelled them out would raise."""
    unknown = sorted(set(kwargs) - set(allowed))
    if unknown:
        raise TypeError(f"unexpected keyword argument(s): {', '.join(unknown)}")


def _check_t2_options(max_terms, optimize: bool, kwargs: dict) -> None:
    check_optimize_kwargs(kwargs, _T2_ARGS)
    if max_terms is not None and max_terms < 1:
        raise ValueError(f"max_terms must be a positive count, got {max_terms}")
    if optimize:
        raise NotImplementedError(
            "optimize=True (compressed double factorization) is not supported; "
            "only the explicit factorization is available."
        )


def double_factorized_t2(t2_amplitudes: np.ndarray, *, tol: float = 1e-8,
                         max_terms: int | None = None, optimize: bool = False,
                         **optimize_kwargs):
    r"""Double factorization :math:`t_{ijab} = i \sum_k \sum_{pq} Z^{(k)}_{pq}
    U^{(k)}_{ap} U^{(k)*}_{ip} U^{(k)}_{bq} U^{(k)*}_{jq}` of real t2 amplitudes
    (python/ffsim/linalg/double_factorized_decomposition.py:467-657).

    The (ia),(jb) matrix of ``t2`` is eigendecomposed (largest |eigenvalue|
    first, ``tol`` dropping the small tail); every kept eigenpair
    :math:`(\lambda, v)` gives two terms, one for each quadrature
    :math:`Q_\pm` of the excitation matrix built from :math:`v`:
    :math:`Z = \pm\lambda\, e e^T` and :math:`U` the eigenvectors of
    :math:`Q_\pm = U\,\mathrm{diag}(e)\,U^\dagger`.  Terms are returned in that
    order, ``(+, -)`` per eigenpair, cut to ``max_terms``.  Returns
    ``(diag_coulomb_mats, orbital_rotations)`` of shape ``(L, norb, norb)``.
    ``optimize=True`` raises ``NotImplementedError``; the keyword arguments
    only the optimized form reads (``method``, ``callback``, ``options``,
    ``diag_coulomb_indices``, ``regularization``, ``multi_stage_start``,
    ``multi_stage_step``, ``return_optimize_result``) are accepted and
    ignored, as ffsim ignores them without it; any other raises ``TypeError``.
    """
    _check_t2_options(max_terms, optimize, optimize_kwargs)
    t2_amplitudes = np.asarray(t2_amplitudes)
    nocc, _, nvrt, _ = t2_amplitudes.shape
    norb = nocc + nvrt
    t2_mat = t2_amplitudes.transpose(0, 2, 1, 3).reshape(nocc * nvrt, nocc * nvrt)
    lam, vecs = _ranked_eigh(t2_mat, tol)
    n_vecs = len(lam)
    quads = np.zeros((n_vecs, 2, norb, norb), dtype=complex)
    for k in range(n_vecs):
        mat = _excitation_matrix(vecs[:, k], nocc, norb)
        quads[k, 0] = _quadrature(mat, 1)
        quads[k, 1] = _quadrature(mat, -1)
    e, rotations = np.linalg.eigh(quads)
    scale = lam[:, None] * np.array([1.0, -1.0])
    mats = scale[:, :, None, None] * e[:, :, :, None] * e[:, :, None, :]
    return (mats.reshape(-1, norb, norb)[:max_terms],
            rotations.reshape(-1, norb, norb)[:max_terms])


def double_factorized_t2_alpha_beta(t2_amplitudes: np.ndarray, *, tol: float = 1e-8,
                                    max_terms: int | None = None, optimize: bool = False,
                                    **optimize_kwargs):
    r"""Double factorization of opposite-spin t2 amplitudes ``t2ab[i, j, a, b]``
    (python/ffsim/linalg/double_factorized_decomposition.py:833-991, 1168-1213).

    The (ia),(jb) matrix is factored by a thin SVD (``tol`` dropping the small
    tail).  Every kept singular triple :math:`(\sigma, l, r)` gives four terms,
    ``(s, m)`` for quadrature sign ``s in (+, -)`` and ``m in (0, 1)``: the
    alpha rotation diagonalizes :math:`Q_s(L)`, the beta one
    :math:`Q_s((-1)^m R)`, and with ``e`` the two eigenvalue lists side by side
    the 2 norb x 2 norb matrix :math:`\tfrac12 c\,\sigma\, e e^T`,
    :math:`c = (1, -1, -1, 1)`, is split into its alpha-alpha, alpha-beta and
    beta-beta blocks.  Returns ``(diag_coulomb_mats, orbital_rotations)`` of
    shape ``(L, 3, norb, norb)`` and ``(L, 2, norb, norb)``.
    """
    _check_t2_options(max_terms, optimize, optimize_kwargs)
    t2_amplitudes = np.asarray(t2_amplitudes)
    nocc_a, nocc_b, nvrt_a, nvrt_b = t2_amplitudes.shape
    norb = nocc_a + nvrt_a
    t2_mat = t2_amplitudes.transpose(0, 2, 1, 3).reshape(nocc_a * nvrt_a, nocc_b * nvrt_b)
    left, sing, right = _ranked_svd(t2_mat, tol)
    n_vecs = len(sing)
    quads = np.zeros((n_vecs, 2, 2, 2, norb, norb), dtype=complex)
    for k in range(n_vecs):
        mat_l = _excitation_matrix(left[:, k], nocc_a, norb)
        mat_r = _excitation_matrix(right[k], nocc_b, norb)
        for si, sign in enumerate((1, -1)):
            quads[k, si, :, 0] = _quadrature(mat_l, sign)
            quads[k, si, 0, 1] = _quadrature(mat_r, sign)
            quads[k, si, 1, 1] = _quadrature(-mat_r, sign)
    quads = quads.reshape(n_vecs, 4, 2, norb, norb)
    e, rotations = np.linalg.eigh(quads)
    e = np.concatenate([e[:, :, 0], e[:, :, 1]], axis=-1)
    scale = 0.5 * np.array([1.0, -1.0, -1.0, 1.0]) * sing[:, None]
    big = scale[:, :, None, None] * e[:, :, :, None] * e[:, :, None, :]
    mats = np.stack([big[:, :, :norb, :norb], big[:, :, :norb, norb:],
                     big[:, :, norb:, norb:]], axis=2)
    return (mats.reshape(-1, 3, norb, norb)[:max_terms],
            rotations.reshape(-1, 2, norb, norb)[:max_terms])


# ---------------------------------------------------------------------------
# Double factorization of a two-body tensor (explicit forms).


def modified_cholesky(mat: np.ndarray, *, tol: float = 1e-8, max_vecs: int | None = None):
    r"""Pivoted (modified) Cholesky vectors of a positive semidefinite matrix,
    :math:`M \approx \sum_i v_i v_i^\dagger`
    (python/ffsim/linalg/double_factorized_decomposition.py:70-131; Motta and
    Zhang, arXiv:1711.02242).  Each step takes the column of the largest
    remaining diagonal error, stopping once it falls below ``tol`` or after
    ``max_vecs`` vectors.  Returns the vectors as the columns of a
    ``(dim, n_vecs)`` array."""
    mat = np.asarray(mat)
    dim = mat.shape[0]
    if not dim:
        return np.empty((0, 0))
    if max_vecs is None:
        max_vecs = dim
    vecs = np.zeros((dim, max_vecs + 1), dtype=mat.dtype)
    err = np.diagonal(mat).real.copy()
    n = 0
    while n <= max_vecs:
        piv = int(np.argmax(err))
        if err[piv] < tol:
            break
        v = mat[:, piv].copy()
        if n:
            v -= vecs[:, :n] @ vecs[piv, :n].conj()
        v /= np.sqrt(err[piv])
        vecs[:, n] = v
        err -= np.abs(v) ** 2
        n += 1
    return vecs[:, : min(n, max_vecs)]


def double_factorized(two_body_tensor: np.ndarray, *, tol: float = 1e-8,
                      max_vecs: int | None = None, optimize: bool = False,
                      cholesky: bool = True, **optimize_kwargs):
    r"""Double factorization of a real two-body tensor,
    :math:`h_{pqrs} = \sum_t \sum_{ij} Z^{(t)}_{ij} U^{(t)}_{pi} U^{(t)*}_{qi}
    U^{(t)}_{rj} U^{(t)*}_{sj}` (python/ffsim/linalg/double_factorized_decomposition.py:145-311).

    The ``(pq), (rs)`` matrix of the tensor is factored into vectors (pivoted
    Cholesky, or with ``cholesky=False`` its eigendecomposition, largest
    |eigenvalue| first); every vector, read as an ``norb x norb`` symmetric
    matrix, is diagonalized, ``U`` being its eigenvectors and ``Z`` the outer
    product of its eigenvalues (times the eigenvalue of the outer
    decomposition in the eigh form).  Returns ``(diag_coulomb_mats,
    orbital_rotations)``, each ``(L, norb, norb)``.  ``optimize=True`` raises
    ``NotImplementedError``; the keyword arguments only the optimized form
    reads are accepted and ignored."""
    check_optimize_kwargs(optimize_kwargs, OPTIMIZE_ARGS | {"diag_coulomb_indices",
                                                            "return_optimize_result"})
    two_body_tensor = np.asarray(two_body_tensor)
    if np.iscomplexobj(two_body_tensor):
        raise ValueError("the two-body tensor must be real: complex tensors cannot be double-factorized here")
    norb = two_body_tensor.shape[0]
    if not norb:
        return np.empty((0, 0, 0)), np.empty((0, 0, 0))
    if optimize:
        raise NotImplementedError(
            "optimize=True (compressed double factorization) is not supported; "
            "only the explicit factorization is available.")
    if max_vecs is None:
        max_vecs = norb * (norb + 1) // 2
    flat = two_body_tensor.reshape(norb * norb, norb * norb)
    if cholesky:
        vecs = modified_cholesky(flat, tol=tol, max_vecs=max_vecs)
        weights = np.ones(vecs.shape[1])
    else:
        weights, vecs = _ranked_eigh(flat, tol)
        weights, vecs = weights[:max_vecs], vecs[:, :max_vecs]
    e, rotations = np.linalg.eigh(vecs.T.reshape(-1, norb, norb))
    mats = weights[:, None, None] * e[:, :, None] * e[:, None, :]
    return mats, rotations
