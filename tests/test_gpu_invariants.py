"""Invariants of the gates on the GPU that need no oracle: the checks of
tests/python/gates/orbital_rotation_test.py:41-258 and
tests/python/gates/num_op_sum_test.py:77-95, restated.

* an orbital rotation equals exp of the one-body operator of log U, applied
  with scipy's expm_multiply through the GPU one-body linop;
* rotating a Slater determinant by the eigenvectors of a Hermitian h gives an
  eigenstate of the one-body operator of h, with the summed eigenvalues;
* rotations compose;
* num-op-sum evolution in the eigenbasis of h is the time evolution of h;
* copy=True leaves the input vector and matrix untouched;
* diagonal Coulomb evolution with one matrix for every spin pair conserves
  <S^2> (tests/python/gates/diag_coulomb_test.py:121-148), with <S^2>
  computed here from S^2 = Sz (Sz + 1) + S- S+ over explicit determinants.

Every test prints the case it runs (flushed) before touching the GPU.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest
import scipy.linalg
import scipy.sparse.linalg

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(4, (2, 1)), (5, (2, 2)), (6, (3, 2)), (6, (0, 3)), (7, (3, 4))]


def _case(*args):
    print("CASE", *args, flush=True)
    sys.stdout.flush()


def _dim(norb, nelec):
    return math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])


def _colex(occ):
    """Address of the string with orbitals ``occ`` occupied: sum_t C(c_t, t + 1)."""
    return sum(math.comb(c, t + 1) for t, c in enumerate(sorted(occ)))


def _hermitian(n, rng):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a + a.T.conj()


def _slater(norb, occ_a, occ_b):
    nb = math.comb(norb, len(occ_b))
    vec = np.zeros(math.comb(norb, len(occ_a)) * nb, dtype=complex)
    vec[_colex(occ_a) * nb + _colex(occ_b)] = 1
    return vec


@pytest.mark.parametrize("norb, nelec", CASES)
def test_orbital_rotation_is_exp_of_one_body(norb, nelec, rng):
    import ffsim_amd

    _case("rotation = expm(one-body log U)", norb, nelec)
    mat = O.random_unitary(norb, rng)
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    got = ffsim_amd.apply_orbital_rotation(vec, mat, norb, nelec)
    op = ffsim_amd.contract.one_body_linop(scipy.linalg.logm(mat), norb, nelec)
    expected = scipy.sparse.linalg.expm_multiply(op, vec, traceA=0)
    np.testing.assert_allclose(got, expected, atol=1e-10, rtol=0)


@pytest.mark.parametrize("seed", range(4))
def test_orbital_rotation_prepares_eigenstates(seed):
    import ffsim_amd

    rng = np.random.default_rng(1000 + seed)
    norb = 6
    n_a, n_b = int(rng.integers(0, norb + 1)), int(rng.integers(0, norb + 1))
    occ_a = sorted(rng.choice(norb, n_a, replace=False).tolist())
    occ_b = sorted(rng.choice(norb, n_b, replace=False).tolist())
    _case("eigenstate", norb, occ_a, occ_b)
    h = _hermitian(norb, rng)
    eigs, vecs = scipy.linalg.eigh(h)
    state = _slater(norb, occ_a, occ_b)
    before = state.copy()
    final = ffsim_amd.apply_orbital_rotation(state, vecs, norb, (n_a, n_b))
    np.testing.assert_allclose(np.linalg.norm(final), 1.0, atol=1e-12)
    got = ffsim_amd.contract.contract_one_body(final, h, norb, (n_a, n_b))
    energy = eigs[occ_a].sum() + eigs[occ_b].sum()
    np.testing.assert_allclose(got, energy * final, atol=1e-10, rtol=0)
    np.testing.assert_array_equal(state, before)


@pytest.mark.parametrize("norb, nelec", CASES)
def test_orbital_rotations_compose(norb, nelec, rng):
    import ffsim_amd

    _case("compose", norb, nelec)
    u1, u2 = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    got = ffsim_amd.apply_orbital_rotation(vec, u1, norb, nelec)
    got = ffsim_amd.apply_orbital_rotation(got, u2 @ u1.T.conj(), norb, nelec)
    expected = ffsim_amd.apply_orbital_rotation(vec, u2, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)


@pytest.mark.parametrize("norb, nelec", CASES)
def test_quadratic_hamiltonian_evolution(norb, nelec, rng):
    import ffsim_amd

    _case("quadratic evolution", norb, nelec)
    h = _hermitian(norb, rng)
    eigs, vecs = scipy.linalg.eigh(h)
    time = float(rng.standard_normal())
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    got = ffsim_amd.apply_num_op_sum_evolution(vec, eigs, time, norb, nelec,
                                               orbital_rotation=vecs)
    op = ffsim_amd.contract.one_body_linop(h, norb, nelec)
    expected = scipy.sparse.linalg.expm_multiply(
        -1j * time * op, vec, traceA=-1j * time * np.sum(np.abs(h)))
    np.testing.assert_allclose(got, expected, atol=1e-10, rtol=0)


def test_no_side_effects_and_large_norm(rng):
    import ffsim_amd

    norb, nelec = 12, (6, 6)
    _case("no side effects / norm", norb, nelec)
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    mat_a, mat_b = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    vec0, a0, b0 = vec.copy(), mat_a.copy(), mat_b.copy()
    out = ffsim_amd.apply_orbital_rotation(vec, (mat_a, mat_b), norb, nelec)
    np.testing.assert_allclose(np.linalg.norm(out), 1.0, atol=1e-12)
    np.testing.assert_array_equal(vec, vec0)
    np.testing.assert_array_equal(mat_a, a0)
    np.testing.assert_array_equal(mat_b, b0)
    j = rng.standard_normal((norb, norb))
    j = j + j.T
    j0 = j.copy()
    out = ffsim_amd.apply_diag_coulomb_evolution(vec, j, 0.7, norb, nelec,
                                                 orbital_rotation=mat_a)
    np.testing.assert_allclose(np.linalg.norm(out), 1.0, atol=1e-12)
    np.testing.assert_array_equal(vec, vec0)
    np.testing.assert_array_equal(j, j0)
    np.testing.assert_array_equal(mat_a, a0)


def spin_square(vec, norb, nelec):
    """<S^2> = Sz (Sz + 1) + |S+ psi|^2 with S+ = sum_p a+_{p alpha} a_{p beta}, on
    determinants a+(alpha string) a+(beta string)|0> (alpha operators first):
    a_{p beta} passes the n_a alpha operators and the beta ones below p, and
    a+_{p alpha} the alpha ones below p."""
    n_a, n_b = nelec
    sz = 0.5 * (n_a - n_b)
    if n_b == 0 or n_a == norb:
        return sz * (sz + 1)
    strs = lambda k: [s for s in range(1 << norb) if bin(s).count("1") == k]  # noqa: E731
    a_src, b_src, a_dst, b_dst = strs(n_a), strs(n_b), strs(n_a + 1), strs(n_b - 1)
    a_rank = {s: i for i, s in enumerate(a_dst)}
    b_rank = {s: i for i, s in enumerate(b_dst)}
    psi = np.asarray(vec).reshape(len(a_src), len(b_src))
    out = np.zeros((len(a_dst), len(b_dst)), dtype=complex)
    for i, sa in enumerate(a_src):
        for j, sb in enumerate(b_src):
            for p in range(norb):
                if (sb >> p) & 1 and not (sa >> p) & 1:
                    below = (1 << p) - 1
                    sign = (-1) ** (n_a + bin(sb & below).count("1") + bin(sa & below).count("1"))
                    out[a_rank[sa | 1 << p], b_rank[sb & ~(1 << p)]] += sign * psi[i, j]
    return sz * (sz + 1) + float(np.vdot(out, out).real)


def test_spin_square_known_values():
    """The helper itself: the two-orbital singlet (|0,1> + |1,0> in the
    alpha-first product basis) has <S^2> = 0, the matching triplet 2, and a
    determinant n_b - n_doubly_occupied + Sz (Sz + 1)."""
    singlet = np.array([0, 1, 1, 0]) / np.sqrt(2)
    triplet = np.array([0, 1, -1, 0]) / np.sqrt(2)
    assert abs(spin_square(singlet, 2, (1, 1))) < 1e-14
    assert abs(spin_square(triplet, 2, (1, 1)) - 2) < 1e-14
    vec = _slater(5, [0, 1, 3], [1, 4])
    assert abs(spin_square(vec, 5, (3, 2)) - (2 - 1 + 0.5 * 1.5)) < 1e-14
    hf = _slater(4, [0, 1], [0, 1])
    assert abs(spin_square(hf, 4, (2, 2))) < 1e-14


@pytest.mark.parametrize("z_rep", [False, True])
@pytest.mark.parametrize("norb, nelec", [(4, (2, 2)), (5, (3, 2)), (5, (1, 3))])
def test_diag_coulomb_evolution_conserves_spin_squared(norb, nelec, z_rep, rng):
    import ffsim_amd

    _case("spin square", norb, nelec, z_rep)
    mat = rng.standard_normal((norb, norb))
    mat = mat @ mat.T
    rot = O.random_unitary(norb, rng)
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    time = float(rng.uniform())
    before = spin_square(vec, norb, nelec)
    out = ffsim_amd.apply_diag_coulomb_evolution(vec, mat, time, norb, nelec,
                                                 orbital_rotation=rot, z_representation=z_rep)
    np.testing.assert_allclose(spin_square(out, norb, nelec), before, rtol=1e-10)


@pytest.mark.parametrize("z_rep", [False, True])
def test_molecular_hamiltonian_on_gpu(z_rep, rng):
    """A molecular Hamiltonian (real (pq|rs) with 8-fold symmetry) converted by
    DoubleFactorizedHamiltonian.from_molecular_hamiltonian: the GPU linear
    operator equals the dense molecular Hamiltonian, and fourth-order Trotter
    steps converge to its exact propagator."""
    from types import SimpleNamespace

    import ffsim_amd

    norb, nelec = 4, (2, 1)
    _case("molecular", norb, nelec, z_rep)
    h1 = rng.standard_normal((norb, norb))
    h1 = 0.3 * (h1 + h1.T)
    lk = rng.standard_normal((5, norb, norb))
    lk = 0.3 * (lk + lk.transpose(0, 2, 1))
    h2 = np.einsum("kpq,krs->pqrs", lk, lk)
    mol = SimpleNamespace(one_body_tensor=h1, two_body_tensor=h2, constant=0.3)
    df = ffsim_amd.DoubleFactorizedHamiltonian.from_molecular_hamiltonian(
        mol, z_representation=z_rep)
    dense = O.molecular_hamiltonian_dense(h1, h2, 0.3, norb, nelec)
    vec = O.random_state_vector(_dim(norb, nelec), rng)
    np.testing.assert_allclose(ffsim_amd.linear_operator(df, norb, nelec) @ vec, dense @ vec,
                               atol=1e-10, rtol=0)
    exact = scipy.linalg.expm(-0.5j * dense) @ vec
    got = ffsim_amd.simulate_trotter_double_factorized(vec, df, 0.5, norb=norb, nelec=nelec,
                                                       n_steps=8, order=2)
    assert np.linalg.norm(got - exact) < 1e-3
