"""Pin the CPU oracle's gates against independent closed forms.

Each restated function of the oracle (ffsim's schedules over C restatements
of its Rust kernels) is checked against mathematics that shares no code with
it: determinant minors (compound matrices), the exponential of a dense
one-body generator, eigenphases from occupation numbers, and dense
Hamiltonians.  The planted-mistake tests show that a plausible error in the
restatement is caught by these pins.
"""

from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"
TOL = 1e-12

CASES = [(1, (1, 0)), (2, (1, 1)), (3, (2, 1)), (3, (1, 2)), (4, (2, 2)), (4, (3, 2)),
         (5, (2, 3)), (5, (0, 5)), (6, (3, 3)), (6, (4, 1))]


def _vec(norb, nelec, rng):
    return O.random_state_vector(math.comb(norb, nelec[0]) * math.comb(norb, nelec[1]), rng)


def _reconstruct(rots, diag, n):
    M = np.diag(diag)
    for c, s, i, j in rots[::-1]:
        G = np.eye(n, dtype=complex)
        G[i, i] = c
        G[i, j] = s
        G[j, i] = -np.conj(s)
        G[j, j] = c
        M = M @ G.conj()
    return M


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_givens_decomposition_reconstructs(impl, rng):
    """tests/python/linalg/givens_test.py:28-41, :74-101 restated."""
    if impl == "oracle":
        decomp = O.givens_decomposition
    else:
        from ffsim_amd.linalg import givens_decomposition as decomp
    for n in (1, 2, 3, 4, 7, 8, 12, 16):
        U = O.random_unitary(n, rng)
        rots, diag = decomp(U)
        assert len(rots) == n * (n - 1) // 2
        for _, _, i, j in rots:
            assert abs(i - j) == 1
        np.testing.assert_allclose(_reconstruct(rots, diag, n), U, atol=1e-12)
    rots, diag = decomp(np.eye(5))
    assert rots == []
    np.testing.assert_allclose(diag, np.ones(5))
    fixture = np.load(GOLDEN / "orbital_rotation-0.npy")
    rots, diag = decomp(fixture)
    np.testing.assert_allclose(_reconstruct(rots, diag, 8), fixture, atol=1e-12)
    with pytest.raises(ValueError):
        decomp(np.zeros((2, 3), dtype=complex))


@pytest.mark.parametrize("norb,nelec", CASES)
def test_orbital_rotation_vs_minors(norb, nelec, rng):
    vec = _vec(norb, nelec, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    for mat, (a, b) in ((Ua, (Ua, Ua)), ((Ua, Ub), (Ua, Ub)), ((None, Ub), (None, Ub)),
                        ((Ua, None), (Ua, None))):
        np.testing.assert_allclose(
            O.apply_orbital_rotation(vec, mat, norb, nelec),
            O.orbital_rotation_closed_form(vec, a, b, norb, nelec),
            atol=TOL,
        )


@pytest.mark.parametrize("norb,nelec", [(3, (2, 1)), (4, (2, 2)), (5, (3, 1))])
def test_orbital_rotation_vs_one_body_exponential(norb, nelec, rng):
    """orbital_rotation_test.py:38-123 restated: apply U == expm(sum log(U)_pq a^dag_p a_q)."""
    vec = _vec(norb, nelec, rng)
    U = O.random_unitary(norb, rng)
    gen = O.one_body_dense(scipy.linalg.logm(U), norb, nelec)
    np.testing.assert_allclose(O.apply_orbital_rotation(vec, U, norb, nelec),
                               scipy.linalg.expm(gen) @ vec, atol=1e-11)


def test_orbital_rotation_special_case_fixture():
    """orbital_rotation_test.py:187-206: fixture unitary on the (5,5) HF state."""
    mat = np.load(GOLDEN / "orbital_rotation-0.npy")
    hf = np.zeros(math.comb(8, 5) ** 2, dtype=complex)
    hf[0] = 1
    result = O.apply_orbital_rotation(hf, mat, 8, (5, 5))
    np.testing.assert_allclose(np.linalg.norm(result), 1, atol=1e-12)
    np.testing.assert_allclose(result, O.orbital_rotation_closed_form(hf, mat, mat, 8, (5, 5)),
                               atol=1e-12)
    gen = O.one_body_dense(scipy.linalg.logm(mat), 8, (5, 5))
    np.testing.assert_allclose(result, scipy.sparse.linalg.expm_multiply(
        scipy.sparse.csr_matrix(gen), hf, traceA=0), atol=1e-12)


def test_orbital_rotation_composition(rng):
    """orbital_rotation_test.py:165-184: rotating by U1 then U2 U1^dag equals U2."""
    norb, nelec = 5, (2, 3)
    vec = _vec(norb, nelec, rng)
    U1, U2 = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    two = O.apply_orbital_rotation(O.apply_orbital_rotation(vec, U1, norb, nelec),
                                   U2 @ U1.T.conj(), norb, nelec)
    np.testing.assert_allclose(two, O.apply_orbital_rotation(vec, U2, norb, nelec), atol=TOL)


@pytest.mark.parametrize("z_rep", [False, True])
@pytest.mark.parametrize("norb,nelec", [(2, (1, 1)), (3, (2, 1)), (4, (2, 2)), (5, (2, 3)),
                                        (6, (3, 3))])
def test_diag_coulomb_eigenphases(norb, nelec, z_rep, rng):
    """diag_coulomb_test.py:149-335 restated: exp(-i t E) on every determinant."""
    vec = _vec(norb, nelec, rng)
    t = 0.8
    J = O.random_symmetric_uniform(norb, rng)
    mats = [J, (J, rng.standard_normal((norb, norb)), O.random_symmetric_uniform(norb, rng)),
            (None, J, None), (J, None, None), (None, None, J)]
    for mat in mats:
        E = O.diag_coulomb_energies(mat, norb, nelec, z_rep)
        np.testing.assert_allclose(
            O.apply_diag_coulomb_evolution(vec, mat, t, norb, nelec, z_representation=z_rep),
            np.exp(-1j * t * E) * vec, atol=TOL)
        np.testing.assert_allclose(
            O.contract_diag_coulomb(vec, mat, norb, nelec, z_representation=z_rep),
            E * vec, atol=TOL)
    U = O.random_unitary(norb, rng)
    E = O.diag_coulomb_energies(J, norb, nelec, z_rep)
    L = np.kron(O.compound_matrix(U, nelec[0]), O.compound_matrix(U, nelec[1]))
    dense = L @ np.diag(np.exp(-1j * t * E)) @ L.conj().T
    np.testing.assert_allclose(
        O.apply_diag_coulomb_evolution(vec, J, t, norb, nelec, orbital_rotation=U,
                                       z_representation=z_rep),
        dense @ vec, atol=TOL)


def test_diag_coulomb_number_rep_vs_diag_formula(rng):
    """The _diag_ formula of diagonal_coulomb_hamiltonian.py:208-238 (numpy occupations)."""
    norb, nelec = 5, (3, 2)
    J_aa, J_ab = O.random_symmetric_uniform(norb, rng), O.random_symmetric_uniform(norb, rng)
    occ_a = O.occupations(O.colex_strings(norb, 3), norb)
    occ_b = O.occupations(O.colex_strings(norb, 2), norb)
    diag = (0.5 * np.sum((occ_a @ J_aa) * occ_a, axis=1)[:, None]
            + 0.5 * np.sum((occ_b @ J_aa) * occ_b, axis=1)[None, :] + occ_a @ J_ab @ occ_b.T)
    vec = _vec(norb, nelec, rng)
    np.testing.assert_allclose(
        O.contract_diag_coulomb(vec, (J_aa, J_ab, J_aa), norb, nelec), diag.reshape(-1) * vec,
        atol=TOL)


@pytest.mark.parametrize("norb,nelec", [(3, (2, 1)), (5, (2, 3)), (6, (3, 3))])
def test_num_op_sum(norb, nelec, rng):
    """num_op_sum_test.py:28-91 restated."""
    vec = _vec(norb, nelec, rng)
    lam, lam_b = rng.standard_normal(norb), rng.standard_normal(norb)
    for coeffs, (ca, cb) in ((lam, (lam, lam)), ((lam, lam_b), (lam, lam_b)),
                             ((None, lam_b), (None, lam_b))):
        E = O.num_op_energies(ca, cb, norb, nelec)
        np.testing.assert_allclose(O.apply_num_op_sum_evolution(vec, coeffs, 0.4, norb, nelec),
                                   np.exp(-0.4j * E) * vec, atol=TOL)
    E = O.num_op_energies(lam, lam, norb, nelec)
    np.testing.assert_allclose(O.contract_num_op_sum(vec, lam, norb, nelec), E * vec, atol=TOL)
    # rotated: U exp(-i t sum lam n) U^dag equals exp(-i t H1) with H1 = U diag(lam) U^dag
    U = O.random_unitary(norb, rng)
    h = U @ np.diag(lam) @ U.conj().T
    np.testing.assert_allclose(
        O.apply_num_op_sum_evolution(vec, lam, 0.4, norb, nelec, orbital_rotation=U),
        scipy.linalg.expm(-0.4j * O.one_body_dense(h, norb, nelec)) @ vec, atol=1e-11)


def _dense_df(df_one_body, mats, rots, constant, norb, nelec, z_rep=False):
    """Dense H_DF from one-body operators and rotated diagonal energies."""
    H = constant * np.eye(math.comb(norb, nelec[0]) * math.comb(norb, nelec[1]), dtype=complex)
    H += O.one_body_dense(df_one_body, norb, nelec)
    for J, U in zip(mats, rots):
        L = np.kron(O.compound_matrix(U, nelec[0]), O.compound_matrix(U, nelec[1]))
        H += L @ np.diag(O.diag_coulomb_energies(J, norb, nelec, z_rep)) @ L.conj().T
    return H


@pytest.mark.parametrize("z_rep", [False, True])
def test_df_hamiltonian_matvec_dense(z_rep, rng):
    norb, nelec = 4, (2, 1)
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = h + h.conj().T
    mats = [O.random_symmetric_uniform(norb, rng) for _ in range(3)]
    rots = [O.random_unitary(norb, rng) for _ in range(3)]
    vec = _vec(norb, nelec, rng)
    got = O.df_hamiltonian_matvec(vec, h, mats, rots, 0.3, norb, nelec, z_rep)
    np.testing.assert_allclose(got, _dense_df(h, mats, rots, 0.3, norb, nelec, z_rep) @ vec,
                               atol=1e-11)


def test_ucj_vs_dense_product(rng):
    norb, nelec, n_reps = 4, (2, 2), 2
    J = np.stack([np.stack([O.random_symmetric_uniform(norb, rng) for _ in range(2)])
                  for _ in range(n_reps)])
    Us = np.stack([O.random_unitary(norb, rng) for _ in range(n_reps)])
    final = O.random_unitary(norb, rng)
    vec = _vec(norb, nelec, rng)
    lam = lambda U: np.kron(O.compound_matrix(U, 2), O.compound_matrix(U, 2))  # noqa: E731
    expected = vec.copy()
    for (j_aa, j_ab), U in zip(J, Us):
        E = O.diag_coulomb_energies((j_aa, j_ab, j_aa), norb, nelec)
        expected = lam(U) @ (np.exp(1j * E) * (lam(U).conj().T @ expected))
    expected = lam(final) @ expected
    np.testing.assert_allclose(O.apply_ucj_spin_balanced(vec, J, Us, final, norb, nelec),
                               expected, atol=TOL)
    np.testing.assert_allclose(
        O.apply_ucj_spin_balanced(vec, np.zeros_like(J), np.stack([np.eye(norb)] * n_reps),
                                  None, norb, nelec), vec, atol=TOL)


def test_golden_outputs_reproduced():
    g = np.load(GOLDEN / "golden.npz")
    norb, nelec, t = 8, (4, 4), float(g["time"])
    rotated = O.apply_orbital_rotation(g["vec"], g["U"], norb, nelec)
    np.testing.assert_allclose(rotated, g["rotated"], atol=1e-13)
    np.testing.assert_allclose(
        O.orbital_rotation_closed_form(g["vec"], g["U"], g["U"], norb, nelec), g["rotated"],
        atol=1e-12)
    E = O.diag_coulomb_energies(g["J"], norb, nelec)
    np.testing.assert_allclose(g["diag_coulomb"], np.exp(-1j * t * E) * g["rotated"], atol=1e-12)
    Ez = O.diag_coulomb_energies(g["J"], norb, nelec, True)
    np.testing.assert_allclose(g["diag_coulomb_z"], np.exp(-1j * t * Ez) * g["rotated"],
                               atol=1e-12)


# ---------------------------------------------------------------- planted mistakes


def test_planted_mistake_in_givens_conjugation_is_caught(rng, monkeypatch):
    """If the oracle applied (c, s) instead of ffsim's (c, conj(s)) the minors pin fails."""
    norb, nelec = 4, (2, 2)
    vec = _vec(norb, nelec, rng)
    U = O.random_unitary(norb, rng)
    real_adjacent = O._apply_adjacent

    def wrong(vec2d, c, s, target_orbs, n, nocc):
        real_adjacent(vec2d, c, np.conj(s), target_orbs, n, nocc)

    monkeypatch.setattr(O, "_apply_adjacent", wrong)
    bad = O.apply_orbital_rotation(vec, U, norb, nelec)
    assert np.abs(bad - O.orbital_rotation_closed_form(vec, U, U, norb, nelec)).max() > 1e-3


def test_planted_mistake_in_diag_halving_is_caught(rng, monkeypatch):
    """Dropping the halved diagonal of _get_mat_exp is caught by the eigenphase pin."""
    norb, nelec = 4, (2, 2)
    vec = _vec(norb, nelec, rng)
    J = O.random_symmetric_uniform(norb, rng)
    real = O.get_mat_exp

    def unhalved(mat, time, n, z):
        m = np.asarray(mat, float)
        e = np.exp(-1j * time * m)
        return e, e, e

    monkeypatch.setattr(O, "get_mat_exp", unhalved)
    bad = O.apply_diag_coulomb_evolution(vec, J, 0.7, norb, nelec)
    E = O.diag_coulomb_energies(J, norb, nelec)
    assert np.abs(bad - np.exp(-0.7j * E) * vec).max() > 1e-3
    monkeypatch.setattr(O, "get_mat_exp", real)


@pytest.mark.parametrize("z_rep", [False, True])
def test_trotter_vs_exact_propagator(z_rep, rng):
    """trotter/double_factorized_test.py:23-83 restated: Trotter converges to
    expm(-i t H) of the dense H_DF; order 1/2 errors shrink as expected."""
    norb, nelec = 4, (2, 1)
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = 0.1 * (h + h.conj().T)
    mats = [0.1 * O.random_symmetric_uniform(norb, rng) for _ in range(3)]
    rots = [O.random_unitary(norb, rng) for _ in range(3)]
    vec = _vec(norb, nelec, rng)
    t = 0.5
    exact = scipy.linalg.expm(-1j * t * _dense_df(h, mats, rots, 0.2, norb, nelec, z_rep)) @ vec
    errs = []
    for order, n_steps in ((0, 10), (1, 10), (2, 3)):
        got = O.simulate_trotter_double_factorized(vec, h, mats, rots, 0.2, t, norb, nelec,
                                                   n_steps=n_steps, order=order,
                                                   z_representation=z_rep)
        errs.append(np.linalg.norm(got - exact))
    assert errs[0] < 3e-2 and errs[1] < 3e-3 and errs[2] < 3e-4, errs


def test_trotter_single_term_is_exact(rng):
    """One diag-Coulomb factor and no one-body term: the formula is exact."""
    norb, nelec = 4, (2, 2)
    J = O.random_symmetric_uniform(norb, rng)
    U = O.random_unitary(norb, rng)
    vec = _vec(norb, nelec, rng)
    got = O.simulate_trotter_double_factorized(vec, np.zeros((norb, norb)), [J], [U], 0.0, 0.9,
                                               norb, nelec, n_steps=1)
    exact = scipy.linalg.expm(-0.9j * _dense_df(np.zeros((norb, norb)), [J], [U], 0.0, norb, nelec)) @ vec
    np.testing.assert_allclose(got, exact, atol=1e-12)


@pytest.mark.parametrize("norb, nelec", [(1, (1, 0)), (3, (2, 1)), (4, (2, 2)), (5, (3, 1)),
                                         (4, (0, 4)), (6, (3, 3))])
def test_contract_one_body_vs_dense(norb, nelec, rng):
    """contract_one_body restated over pyscf's link tables vs the dense one-body
    matrix built from creation/annihilation signs."""
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vec = O.random_state_vector(dim, rng)
    mat = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    expected = O.one_body_dense(mat, norb, nelec) @ vec
    np.testing.assert_allclose(O.contract_one_body(vec, mat, norb, nelec), expected, atol=TOL)


def test_linkstr_index_known_answer():
    """pyscf's gen_linkstr_index layout for norb=3, one electron: strings 0b001,
    0b010, 0b100; (i, i) first, then i -> a with the sign of a+_a a_i."""
    links = O.gen_linkstr_index(3, 1)
    np.testing.assert_array_equal(links[0], [[0, 0, 0, 1], [1, 0, 1, 1], [2, 0, 2, 1]])
    np.testing.assert_array_equal(links[1], [[1, 1, 1, 1], [0, 1, 0, 1], [2, 1, 2, 1]])
    # two electrons: moving orbital 0 to 2 across an occupied orbital 1 flips the sign
    links2 = O.gen_linkstr_index(3, 2)
    assert [2, 0, 2, -1] in links2[0].tolist()  # 0b011 -> 0b110, address 2


def _dense_dc(h, aa, ab, constant, norb, nelec):
    """Dense diagonal Coulomb Hamiltonian: one-body part plus diagonal energies
    (J_aa = J_bb = aa, J_ab = ab)."""
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    return (constant * np.eye(dim) + O.one_body_dense(h, norb, nelec)
            + np.diag(O.diag_coulomb_energies((aa, ab, aa), norb, nelec)))


def test_split_op_trotter_vs_exact_propagator(rng):
    """tests/python/trotter/diag_coulomb_split_op_test.py restated: the split-operator
    formula converges to expm(-i t H) of the dense Hamiltonian, with the errors
    of orders 0, 1 and 2 shrinking as expected; each half alone is exact."""
    norb, nelec = 4, (2, 1)
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = 0.2 * (h + h.conj().T)
    aa = 0.2 * O.random_symmetric_uniform(norb, rng)
    ab = 0.2 * O.random_symmetric_uniform(norb, rng)
    vec = _vec(norb, nelec, rng)
    t = 0.7
    exact = scipy.linalg.expm(-1j * t * _dense_dc(h, aa, ab, 0.3, norb, nelec)) @ vec
    errs = []
    for order, n_steps in ((0, 10), (1, 10), (2, 3)):
        got = O.simulate_trotter_diag_coulomb_split_op(vec, h, (aa, ab), 0.3, t, norb, nelec,
                                                       n_steps=n_steps, order=order)
        errs.append(np.linalg.norm(got - exact))
    assert errs[0] < 3e-2 and errs[1] < 3e-3 and errs[2] < 3e-4, errs
    zero = np.zeros((norb, norb))
    for hh, a, b in ((h, zero, zero), (zero, aa, ab)):
        got = O.simulate_trotter_diag_coulomb_split_op(vec, hh, (a, b), 0.0, t, norb, nelec)
        want = scipy.linalg.expm(-1j * t * _dense_dc(hh, a, b, 0.0, norb, nelec)) @ vec
        np.testing.assert_allclose(got, want, atol=1e-12)


@pytest.mark.parametrize("norb, nelec", [(4, (2, 1)), (5, (2, 2))])
def test_df_z_representation_is_the_same_operator(norb, nelec, rng):
    """DoubleFactorizedHamiltonian.to_z_representation / to_number_representation
    (double_factorized_hamiltonian.py:107-138, :308-330): the dense operator built
    from the number-representation energies equals the one built from the
    Z-representation energies after the conversion, and the round trip is exact."""
    from ffsim_amd.hamiltonians import DoubleFactorizedHamiltonian

    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = h + h.conj().T
    mats = np.stack([O.random_symmetric_uniform(norb, rng) for _ in range(3)])
    rots = np.stack([O.random_unitary(norb, rng) for _ in range(3)])
    ham = DoubleFactorizedHamiltonian(h, mats, rots, constant=0.4)
    z = ham.to_z_representation()
    assert z.z_representation and z.to_z_representation() is z
    np.testing.assert_allclose(
        _dense_df(z.one_body_tensor, mats, rots, z.constant, norb, nelec, z_rep=True),
        _dense_df(h, mats, rots, 0.4, norb, nelec, z_rep=False), atol=1e-11)
    back = z.to_number_representation()
    assert not back.z_representation and ham.to_number_representation() is ham
    assert back._approx_eq_(ham, rtol=0, atol=1e-12)
    assert not z._approx_eq_(ham, rtol=1e-5, atol=1e-8)


@pytest.mark.parametrize("z_rep", [False, True])
@pytest.mark.parametrize("norb, nelec", [(4, (2, 1)), (5, (2, 2)), (3, (0, 3))])
def test_df_hamiltonian_diag_vs_dense(norb, nelec, z_rep, rng):
    """DoubleFactorizedHamiltonian._diag_ (double_factorized_hamiltonian.py:277-280)
    equals the diagonal of the dense operator built from rotated eigenphase
    energies."""
    from ffsim_amd.hamiltonians import DoubleFactorizedHamiltonian

    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = h + h.conj().T
    mats = np.stack([O.random_symmetric_uniform(norb, rng) for _ in range(3)])
    rots = np.stack([O.random_unitary(norb, rng) for _ in range(3)])
    ham = DoubleFactorizedHamiltonian(h, mats, rots, constant=0.7, z_representation=z_rep)
    dense = _dense_df(h, mats, rots, 0.7, norb, nelec, z_rep=z_rep)
    np.testing.assert_allclose(ham._diag_(norb, nelec), np.diag(dense).real, atol=1e-11)


def _random_two_body(norb, rng, rank=5):
    """Real (pq|rs) with the 8-fold symmetry and a positive semidefinite
    (pq),(rs) matrix: sum_k L_k (x) L_k with symmetric L_k."""
    L = rng.standard_normal((rank, norb, norb))
    L = L + L.transpose(0, 2, 1)
    return np.einsum("kpq,krs->pqrs", L, L)


@pytest.mark.parametrize("cholesky", [True, False])
@pytest.mark.parametrize("z_rep", [False, True])
def test_df_from_molecular_hamiltonian_vs_dense(cholesky, z_rep, rng):
    """DoubleFactorizedHamiltonian.from_molecular_hamiltonian
    (double_factorized_hamiltonian.py:140-222) with linalg.double_factorized
    (double_factorized_decomposition.py:145-311): the dense double-factorized
    operator equals the dense molecular Hamiltonian; the factors rebuild the
    two-body tensor."""
    from types import SimpleNamespace

    from ffsim_amd import linalg
    from ffsim_amd.hamiltonians import DoubleFactorizedHamiltonian

    norb, nelec = 4, (2, 1)
    h1 = rng.standard_normal((norb, norb))
    h1 = h1 + h1.T
    h2 = _random_two_body(norb, rng)
    mol = SimpleNamespace(one_body_tensor=h1, two_body_tensor=h2, constant=0.3)
    df = DoubleFactorizedHamiltonian.from_molecular_hamiltonian(mol, cholesky=cholesky,
                                                                z_representation=z_rep)
    assert df.z_representation == z_rep
    dense = _dense_df(df.one_body_tensor, df.diag_coulomb_mats, df.orbital_rotations,
                      df.constant, norb, nelec, z_rep=z_rep)
    np.testing.assert_allclose(dense, O.molecular_hamiltonian_dense(h1, h2, 0.3, norb, nelec),
                               atol=1e-10)
    mats, rots = linalg.double_factorized(h2, cholesky=cholesky)
    rebuilt = np.einsum("kij,kpi,kqi,krj,ksj->pqrs", mats, rots, rots.conj(), rots, rots.conj())
    np.testing.assert_allclose(rebuilt, h2, atol=1e-10)


def test_modified_cholesky_and_truncation(rng):
    """modified_cholesky (double_factorized_decomposition.py:70-131): exact for a
    positive semidefinite matrix, truncated by max_vecs, stopped by tol."""
    from ffsim_amd import linalg

    a = rng.standard_normal((8, 5)) + 1j * rng.standard_normal((8, 5))
    m = a @ a.conj().T  # rank 5
    v = linalg.modified_cholesky(m)
    assert v.shape == (8, 5)
    np.testing.assert_allclose(v @ v.conj().T, m, atol=1e-10)
    assert linalg.modified_cholesky(m, max_vecs=3).shape == (8, 3)
    assert linalg.modified_cholesky(np.zeros((3, 3))).shape == (3, 0)
    h2 = _random_two_body(4, rng, rank=6)
    mats, _ = linalg.double_factorized(h2, max_vecs=2)
    assert len(mats) == 2
    with pytest.raises(ValueError):
        linalg.double_factorized(h2 + 0j)
    with pytest.raises(NotImplementedError):
        linalg.double_factorized(h2, optimize=True)
