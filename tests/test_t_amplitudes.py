"""UCJ operators from t amplitudes: the explicit double factorizations of t2
(python/ffsim/linalg/double_factorized_decomposition.py:467-657, 833-1213) and
the constructors built on them (variational/ucj_spin_balanced.py:377-540,
ucj_spinless.py:322-453, ucj_spin_unbalanced.py:430-703,
ucj_angles_spin_balanced.py:143-158).

The factorizations are pinned by what fixes them mathematically: the terms
reconstruct t2 through the defining sum
t_ijab = i sum_k sum_pq Z_pq U_ap U*_ip U_bq U*_jq (written here with explicit
loops over the terms), the first 2k terms reproduce the top-k eigen-sum of the
(ia),(jb) matrix, the dropped tail stays within ``tol``, and the pair relations
between terms that ffsim's own tests assert hold.  ffsim's energy known answer
(N2/sto-6g, -108.592087) needs pyscf CCSD amplitudes, which are not available
here.
"""

from __future__ import annotations

import itertools

import numpy as np
import pytest
import scipy.linalg

from ffsim_amd import linalg
from ffsim_amd.random import random_t2_amplitudes
from ffsim_amd.variational import (
    UCJAnglesOpSpinBalanced,
    UCJOpSpinBalanced,
    UCJOpSpinless,
    UCJOpSpinUnbalanced,
)
from ffsim_amd.variational.util import orbital_rotation_from_t1_amplitudes

RNG = np.random.default_rng(8123)


def _rebuild(mats, rots):
    """Full norb^4 tensor i sum_k sum_pq Z_pq U_ap U*_ip U_bq U*_jq, term by term."""
    norb = rots.shape[-1]
    out = np.zeros((norb,) * 4, dtype=complex)
    for z, u in zip(mats, rots):
        w = u * 1.0
        # x[a, i, p] = U_ap U*_ip
        x = w[:, None, :] * w.conj()[None, :, :]
        out += 1j * np.einsum("pq,aip,bjq->ijab", z, x, x)
    return out


def _rebuild_t2(mats, rots, nocc):
    return _rebuild(mats, rots)[:nocc, :nocc, nocc:, nocc:]


def _rebuild_t2ab(mats, rots, norb, nocc_a, nocc_b):
    big_z = np.array([np.block([[m[0], m[1]], [m[1].T, m[2]]]) for m in mats])
    big_u = np.array([scipy.linalg.block_diag(u[0], u[1]) for u in rots])
    full = _rebuild(big_z, big_u)
    return full[:nocc_a, norb:norb + nocc_b, nocc_a:norb, norb + nocc_b:]


def _t2_matrix(t2):
    nocc, _, nvrt, _ = t2.shape
    return t2.transpose(0, 2, 1, 3).reshape(nocc * nvrt, nocc * nvrt)


@pytest.mark.parametrize("norb, nocc", [(4, 2), (5, 2), (5, 3), (6, 1)])
def test_df_t2_reconstructs(norb, nocc):
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    mats, rots = linalg.double_factorized_t2(t2)
    np.testing.assert_allclose(_rebuild_t2(mats, rots, nocc), t2, atol=1e-8)
    assert mats.dtype == np.float64 and rots.shape == mats.shape
    for u in rots:
        np.testing.assert_allclose(u @ u.conj().T, np.eye(norb), atol=1e-12)
    for z in mats:
        np.testing.assert_allclose(z, z.T, atol=1e-14)
    np.testing.assert_allclose(mats[0::2], -mats[1::2], atol=1e-8)
    np.testing.assert_allclose(rots[0::2], rots[1::2].conj(), atol=1e-8)


def test_df_t2_terms_follow_ranked_eigenpairs():
    norb, nocc = 6, 2
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    lam, vec = np.linalg.eigh(_t2_matrix(t2))
    order = np.argsort(-np.abs(lam))
    for k in (1, 2, 3):
        mats, rots = linalg.double_factorized_t2(t2, max_terms=2 * k)
        assert len(mats) == 2 * k
        top = sum(lam[m] * np.outer(vec[:, m], vec[:, m]) for m in order[:k])
        got = _t2_matrix(_rebuild_t2(mats, rots, nocc))
        np.testing.assert_allclose(got, top, atol=1e-10)
    mats, _ = linalg.double_factorized_t2(t2, max_terms=3)
    assert len(mats) == 3


def test_df_t2_tol_bounds_error():
    norb, nocc = 6, 2
    nvrt = norb - nocc
    n = nocc * nvrt
    q = np.linalg.qr(RNG.standard_normal((n, n)))[0]
    lam = np.array([3.0, -2.0, 1.0, 1e-4, -2e-4, 1e-5, 0, 0])
    mat = (q * lam) @ q.T
    t2 = mat.reshape(nocc, nvrt, nocc, nvrt).transpose(0, 2, 1, 3)
    mats, rots = linalg.double_factorized_t2(t2, tol=1e-3)
    assert len(mats) == 6
    err = np.max(np.abs(_rebuild_t2(mats, rots, nocc) - t2))
    assert err <= 1e-3
    mats, _ = linalg.double_factorized_t2(t2, tol=1e-3, max_terms=4)
    assert len(mats) == 4


def test_df_t2_errors():
    t2 = random_t2_amplitudes(4, 2, seed=RNG, dtype=float)
    with pytest.raises(ValueError, match="max_terms"):
        linalg.double_factorized_t2(t2, max_terms=0)
    with pytest.raises(NotImplementedError):
        linalg.double_factorized_t2(t2, optimize=True)
    with pytest.raises(ValueError, match="max_terms"):
        linalg.double_factorized_t2_alpha_beta(RNG.standard_normal((2, 2, 3, 3)), max_terms=0)


@pytest.mark.parametrize("shape", [(3, 6, 7, 4), (2, 2, 3, 3), (1, 3, 4, 2)])
def test_df_t2_alpha_beta_reconstructs(shape):
    t2ab = RNG.standard_normal(shape)
    nocc_a, nocc_b, nvrt_a, _ = shape
    norb = nocc_a + nvrt_a
    mats, rots = linalg.double_factorized_t2_alpha_beta(t2ab)
    assert mats.shape[1:] == (3, norb, norb) and rots.shape[1:] == (2, norb, norb)
    np.testing.assert_allclose(_rebuild_t2ab(mats, rots, norb, nocc_a, nocc_b), t2ab,
                               atol=1e-8)
    np.testing.assert_allclose(mats[0::4], -mats[1::4], atol=1e-8)
    np.testing.assert_allclose(mats[0::4], -mats[2::4], atol=1e-8)
    np.testing.assert_allclose(mats[0::4], mats[3::4], atol=1e-8)
    np.testing.assert_allclose(rots[0::4, 0], rots[1::4, 0], atol=1e-8)
    np.testing.assert_allclose(rots[0::4, 0], rots[2::4, 0].conj(), atol=1e-8)
    np.testing.assert_allclose(rots[0::4, 0], rots[3::4, 0].conj(), atol=1e-8)


def test_df_t2_alpha_beta_truncation_follows_singular_values():
    shape = (2, 3, 4, 3)
    t2ab = RNG.standard_normal(shape)
    nocc_a, nocc_b, nvrt_a, nvrt_b = shape
    norb = nocc_a + nvrt_a
    u, s, vh = np.linalg.svd(t2ab.transpose(0, 2, 1, 3).reshape(nocc_a * nvrt_a, -1))
    for k in (1, 2):
        mats, rots = linalg.double_factorized_t2_alpha_beta(t2ab, max_terms=4 * k)
        got = _rebuild_t2ab(mats, rots, norb, nocc_a, nocc_b)
        top = (u[:, :k] * s[:k]) @ vh[:k]
        np.testing.assert_allclose(got.transpose(0, 2, 1, 3).reshape(top.shape), top,
                                   atol=1e-10)


def test_random_t2_symmetry_and_draw_order():
    norb, nocc = 5, 2
    t2 = random_t2_amplitudes(norb, nocc, seed=11, dtype=float)
    np.testing.assert_array_equal(t2, t2.transpose(1, 0, 3, 2))
    nvrt = norb - nocc
    exc = list(itertools.product(range(nocc), range(nvrt)))
    draws = np.random.default_rng(11).standard_normal(len(exc) * (len(exc) + 1) // 2)
    got = [t2[i, j, a, b] for (i, a), (j, b) in itertools.combinations_with_replacement(exc, 2)]
    np.testing.assert_array_equal(got, draws)
    c = random_t2_amplitudes(norb, nocc, seed=11)
    assert np.iscomplexobj(c)
    np.testing.assert_allclose(c, c.transpose(1, 0, 3, 2), atol=1e-12)


def test_t1_rotation():
    nocc, nvrt = 3, 4
    norb = nocc + nvrt
    t1 = RNG.standard_normal((nocc, nvrt)) + 1j * RNG.standard_normal((nocc, nvrt))
    u = orbital_rotation_from_t1_amplitudes(t1)
    gen = np.zeros((norb, norb), dtype=complex)
    for i in range(nocc):
        for a in range(nvrt):
            gen[nocc + a, i] = t1[i, a]
            gen[i, nocc + a] = -np.conj(t1[i, a])
    w, v = np.linalg.eigh(1j * gen)  # 1j*gen is Hermitian
    np.testing.assert_allclose(u, (v * np.exp(-1j * w)) @ v.conj().T, atol=1e-12)
    np.testing.assert_allclose(u @ u.conj().T, np.eye(norb), atol=1e-12)


@pytest.mark.parametrize("n_reps", [None, 3, 15])
def test_ucj_spin_balanced_from_t_amplitudes(n_reps):
    norb, nocc = 5, 3
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    t1 = RNG.standard_normal((nocc, norb - nocc))
    op = UCJOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=n_reps)
    mats, rots = linalg.double_factorized_t2(t2)
    n_terms = len(mats)
    assert op.n_reps == (n_terms if n_reps is None else n_reps)
    assert len(op.to_parameters()) == UCJOpSpinBalanced.n_params(
        norb, op.n_reps, with_final_orbital_rotation=True)
    k = min(op.n_reps, n_terms)
    np.testing.assert_array_equal(op.diag_coulomb_mats[:k, 0], mats[:k])
    np.testing.assert_array_equal(op.diag_coulomb_mats[:k, 1], mats[:k])
    np.testing.assert_array_equal(op.orbital_rotations[:k], rots[:k])
    np.testing.assert_array_equal(op.diag_coulomb_mats[k:], 0)
    for u in op.orbital_rotations[k:]:
        np.testing.assert_array_equal(u, np.eye(norb))
    np.testing.assert_array_equal(op.final_orbital_rotation,
                                  orbital_rotation_from_t1_amplitudes(t1))
    if op.n_reps >= n_terms:
        np.testing.assert_allclose(
            _rebuild_t2(op.diag_coulomb_mats[:, 0], op.orbital_rotations, nocc), t2,
            atol=1e-8)


def test_ucj_spin_balanced_from_zero_t_amplitudes():
    norb, nocc = 5, 3
    for n_reps in (3, 15):
        op = UCJOpSpinBalanced.from_t_amplitudes(
            np.zeros((nocc, nocc, norb - nocc, norb - nocc)),
            t1=np.zeros((nocc, norb - nocc)), n_reps=n_reps)
        assert op.n_reps == n_reps
        np.testing.assert_array_equal(op.diag_coulomb_mats, 0)
        np.testing.assert_allclose(op.orbital_rotations, np.broadcast_to(np.eye(norb),
                                                                         (n_reps, norb, norb)))
        np.testing.assert_allclose(op.final_orbital_rotation, np.eye(norb))
    with pytest.raises(ValueError, match="n_reps"):
        UCJOpSpinBalanced.from_t_amplitudes(np.zeros((2, 2, 2, 2)), n_reps=0)


@pytest.mark.parametrize("pairs_aa, pairs_ab", [
    ("chain", "diag"), ("chain", []), ([], "diag"), ([], []), (None, "diag")])
def test_ucj_spin_balanced_restrict_indices(pairs_aa, pairs_ab):
    norb, nocc = 4, 2
    chain = [(p, p + 1) for p in range(norb - 1)]
    diag = [(p, p) for p in range(norb)]
    pairs_aa = chain if pairs_aa == "chain" else pairs_aa
    pairs_ab = diag if pairs_ab == "diag" else pairs_ab
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    op = UCJOpSpinBalanced.from_t_amplitudes(t2, n_reps=2, interaction_pairs=(pairs_aa, pairs_ab))
    back = UCJOpSpinBalanced.from_parameters(
        op.to_parameters(interaction_pairs=(pairs_aa, pairs_ab)), norb=norb, n_reps=2,
        interaction_pairs=(pairs_aa, pairs_ab))
    np.testing.assert_allclose(back.diag_coulomb_mats, op.diag_coulomb_mats, atol=1e-12)
    np.testing.assert_allclose(back.orbital_rotations, op.orbital_rotations, atol=1e-10)
    with pytest.raises(ValueError, match="lower triangular"):
        UCJOpSpinBalanced.from_t_amplitudes(t2, interaction_pairs=([(1, 0)], None))


def test_ucj_spinless_from_t_amplitudes():
    norb, nocc = 6, 2
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    t1 = RNG.standard_normal((nocc, norb - nocc))
    mats, rots = linalg.double_factorized_t2(t2)
    op = UCJOpSpinless.from_t_amplitudes(t2, t1=t1)
    np.testing.assert_array_equal(op.diag_coulomb_mats, mats)
    np.testing.assert_array_equal(op.orbital_rotations, rots)
    op = UCJOpSpinless.from_t_amplitudes(t2, t1=t1, n_reps=len(mats) + 2)
    assert op.n_reps == len(mats) + 2
    assert len(op.to_parameters()) == UCJOpSpinless.n_params(
        norb, op.n_reps, with_final_orbital_rotation=True)
    chain = [(p, p + 1) for p in range(norb - 1)]
    op = UCJOpSpinless.from_t_amplitudes(t2, n_reps=2, interaction_pairs=chain)
    mask = np.zeros((norb, norb), dtype=bool)
    for p, q in chain:
        mask[p, q] = mask[q, p] = True
    np.testing.assert_array_equal(op.diag_coulomb_mats[:, ~mask], 0)
    np.testing.assert_array_equal(op.diag_coulomb_mats[:, mask], mats[:2][:, mask])


def _t2_spin(norb, nocc_a, nocc_b):
    t2aa = random_t2_amplitudes(norb, nocc_a, seed=RNG, dtype=float)
    t2aa = t2aa - t2aa.transpose(1, 0, 2, 3)
    t2bb = random_t2_amplitudes(norb, nocc_b, seed=RNG, dtype=float)
    t2bb = t2bb - t2bb.transpose(1, 0, 2, 3)
    t2ab = RNG.standard_normal((nocc_a, nocc_b, norb - nocc_a, norb - nocc_b))
    return t2aa, t2ab, t2bb


def test_ucj_spin_unbalanced_from_t_amplitudes():
    norb, nocc_a, nocc_b = 5, 3, 2
    t2aa, t2ab, t2bb = _t2_spin(norb, nocc_a, nocc_b)
    t1 = (RNG.standard_normal((nocc_a, norb - nocc_a)),
          RNG.standard_normal((nocc_b, norb - nocc_b)))
    m_ab, r_ab = linalg.double_factorized_t2_alpha_beta(t2ab)
    m_aa, r_aa = linalg.double_factorized_t2(t2aa)
    m_bb, r_bb = linalg.double_factorized_t2(t2bb)
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), t1=t1)
    n_ab, n_same = len(m_ab), max(len(m_aa), len(m_bb))
    assert op.n_reps == n_ab + n_same
    np.testing.assert_array_equal(op.diag_coulomb_mats[:n_ab], m_ab)
    np.testing.assert_array_equal(op.orbital_rotations[:n_ab], r_ab)
    same = op.diag_coulomb_mats[n_ab:]
    np.testing.assert_array_equal(same[: len(m_aa), 0], m_aa)
    np.testing.assert_array_equal(same[: len(m_bb), 2], m_bb)
    np.testing.assert_array_equal(same[:, 1], 0)
    # The terms rebuild every spin block of t2.
    np.testing.assert_allclose(_rebuild_t2ab(op.diag_coulomb_mats[:n_ab],
                                             op.orbital_rotations[:n_ab], norb, nocc_a, nocc_b),
                               t2ab, atol=1e-8)
    srot = op.orbital_rotations[n_ab:]
    np.testing.assert_allclose(_rebuild_t2(same[:, 0], srot[:, 0], nocc_a), t2aa, atol=1e-8)
    np.testing.assert_allclose(_rebuild_t2(same[:, 2], srot[:, 1], nocc_b), t2bb, atol=1e-8)
    np.testing.assert_array_equal(op.final_orbital_rotation[1],
                                  orbital_rotation_from_t1_amplitudes(t1[1]))
    # Caps: an int caps the total (alpha-beta first), a pair caps each kind.
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=n_ab + 1)
    assert op.n_reps == n_ab + 1
    np.testing.assert_array_equal(op.diag_coulomb_mats[n_ab, 0], m_aa[0])
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=3)
    np.testing.assert_array_equal(op.diag_coulomb_mats, m_ab[:3])
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=(2, 3))
    assert op.n_reps == 5
    np.testing.assert_array_equal(op.diag_coulomb_mats[:2], m_ab[:2])
    np.testing.assert_array_equal(op.diag_coulomb_mats[2:, 0], m_aa[:3])
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=(0, 2))
    assert op.n_reps == 2
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=(1, 40))
    assert op.n_reps == 41
    np.testing.assert_array_equal(op.diag_coulomb_mats[1 + n_same:], 0)
    assert len(op.to_parameters()) == UCJOpSpinUnbalanced.n_params(norb, 41)
    with pytest.raises(ValueError, match="alpha-beta"):
        UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=(-1, 2))
    pairs = ([(0, 1)], [(1, 0), (2, 2)], [])
    op = UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), n_reps=4,
                                               interaction_pairs=pairs)
    keep_ab = np.zeros((norb, norb), dtype=bool)
    keep_ab[1, 0] = keep_ab[2, 2] = True
    np.testing.assert_array_equal(op.diag_coulomb_mats[:, 1, ~keep_ab], 0)
    np.testing.assert_array_equal(op.diag_coulomb_mats[:, 2], 0)


def test_ucj_angles_from_t_amplitudes_matches_matrix_form():
    norb, nocc = 4, 2
    t2 = random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    t1 = RNG.standard_normal((nocc, norb - nocc))
    op = UCJAnglesOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=3)
    ref = UCJAnglesOpSpinBalanced.from_ucj_op(
        UCJOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=3))
    np.testing.assert_array_equal(op.to_parameters(), ref.to_parameters())


def test_ucj_from_cisd_vec_matches_t_amplitudes():
    """ucj_spin_balanced_test.py:288-316 with a synthetic CISD vector laid out as
    the reference documents (c0, singles, doubles; ucj_spin_balanced.py:566-568);
    the conversion to t amplitudes is written out element by element."""
    norb, nocc = 5, 2
    nvrt = norb - nocc
    c0 = 0.9
    c1 = 0.1 * RNG.standard_normal((nocc, nvrt))
    c2 = 0.1 * random_t2_amplitudes(norb, nocc, seed=RNG, dtype=float)
    vec = np.concatenate([[c0], c1.ravel(), c2.ravel()])
    before = vec.copy()
    t1 = c1 / c0
    t2 = np.zeros_like(c2)
    for i, j, a, b in itertools.product(range(nocc), range(nocc), range(nvrt), range(nvrt)):
        t2[i, j, a, b] = c2[i, j, a, b] / c0 - 0.5 * t1[i, a] * t1[j, b]
    op = UCJOpSpinBalanced.from_cisd_vec(vec, norb=norb, nocc=nocc, n_reps=4)
    ref = UCJOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=4)
    np.testing.assert_allclose(op.diag_coulomb_mats, ref.diag_coulomb_mats, atol=1e-12)
    np.testing.assert_allclose(op.orbital_rotations, ref.orbital_rotations, atol=1e-12)
    np.testing.assert_allclose(op.final_orbital_rotation, ref.final_orbital_rotation,
                               atol=1e-12)
    np.testing.assert_array_equal(vec, before)
    with pytest.raises(ValueError, match="c0"):
        UCJOpSpinBalanced.from_cisd_vec(np.concatenate([[0.0], vec[1:]]), norb=norb, nocc=nocc)
    with pytest.raises(ValueError, match="entries"):
        UCJOpSpinBalanced.from_cisd_vec(vec[:-1], norb=norb, nocc=nocc)


def test_optimize_only_arguments_are_accepted_and_others_rejected():
    t2 = random_t2_amplitudes(4, 2, seed=RNG, dtype=float)
    a = linalg.double_factorized_t2(t2, method="BFGS", options={"maxiter": 3},
                                    diag_coulomb_indices=[(0, 0)], return_optimize_result=True)
    b = linalg.double_factorized_t2(t2)
    np.testing.assert_array_equal(a[0], b[0])
    UCJOpSpinBalanced.from_t_amplitudes(t2, n_reps=2, regularization=0.1, multi_stage_step=1)
    with pytest.raises(TypeError, match="metod"):
        UCJOpSpinBalanced.from_t_amplitudes(t2, metod="BFGS")
    with pytest.raises(TypeError, match="diag_coulomb_indices"):
        UCJOpSpinless.from_t_amplitudes(t2, diag_coulomb_indices=[(0, 0)])
    with pytest.raises(TypeError, match="tolerance"):
        linalg.double_factorized_t2_alpha_beta(RNG.standard_normal((2, 2, 2, 2)), tolerance=1)
