"""Adjoint gradient: host chain rule and the oracle's reductions, on the CPU.

The chain rule of ``ffsim_amd.variational.gradient`` (expm adjoint, products
V_k = U_k^dag U_{k-1}, symmetric pair parameters) runs here over primitive
steps supplied by the CPU oracle, and is checked against the oracle's central
finite differences (the reference's gradient mechanism,
python/ffsim/optimize/_util.py:60-110).  The oracle's transition RDM and
phase reductions are pinned against dense one-body matrices, the derivative
of the minors closed form and the derivative of the eigenphase closed form.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.linalg

from oracle import oracle as O
from ffsim_amd.variational.gradient import UCJLayout, _expm_adjoint, ucj_adjoint_sweep
from oracle_adjoint import OracleAdjointOps, five_point_gradient


def _random_hermitian(n, rng):
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return a + a.conj().T


def _dim(norb, nelec):
    return math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])


@pytest.mark.parametrize("norb, nelec", [(3, (2, 1)), (4, (2, 2)), (5, (3, 2))])
def test_transition_rdm_vs_dense_one_body(norb, nelec, rng):
    dim = _dim(norb, nelec)
    lam, psi = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    rdm = O.transition_rdm(lam, psi, norb, nelec)
    X = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    # <lam| sum_pq X_pq (a+_pa a_qa + a+_pb a_qb) |psi> from the dense matrices
    expected = np.vdot(lam, O.one_body_dense(X, norb, nelec) @ psi)
    got = np.sum(X * (rdm[0] + rdm[1]))
    assert abs(got - expected) < 1e-12
    # spin resolution: an alpha-only operator sees only rdm[0]
    na, nb = nelec
    sa = O.colex_strings(norb, na)
    Ha = sum(X[p, q] * O._apply_one_body_term(sa, norb, p, q)
             for p in range(norb) for q in range(norb))
    expected_a = np.vdot(lam, np.kron(Ha, np.eye(math.comb(norb, nb))) @ psi)
    assert abs(np.sum(X * rdm[0]) - expected_a) < 1e-12


def test_transition_rdm_is_derivative_of_minors(rng):
    """d/de <lam| Lambda(e^{eX}) psi> at e=0 equals sum_pq X_pq rho_pq."""
    norb, nelec = 4, (2, 1)
    dim = _dim(norb, nelec)
    lam, psi = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    A = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    X = A - A.conj().T
    rdm = O.transition_rdm(lam, psi, norb, nelec).sum(axis=0)
    h = 1e-6

    def f(e):
        U = scipy.linalg.expm(e * X)
        return np.vdot(lam, O.orbital_rotation_closed_form(psi, U, U, norb, nelec))

    fd = (f(h) - f(-h)) / (2 * h)
    assert abs(fd - np.sum(X * rdm)) < 1e-8


def test_transition_rdm_planted_sign_mistake_is_caught(rng):
    norb, nelec = 4, (2, 2)
    dim = _dim(norb, nelec)
    lam, psi = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    rdm = O.transition_rdm(lam, psi, norb, nelec)
    X = rng.standard_normal((norb, norb))
    expected = np.vdot(lam, O.one_body_dense(X, norb, nelec) @ psi)
    bad = rdm.copy()
    bad[0] = bad[0].T  # a+_q a_p instead of a+_p a_q
    assert abs(np.sum(X * (bad[0] + bad[1])) - expected) > 1e-6


def test_phase_reductions_are_eigenphase_derivatives(rng):
    """d/dJ Re<lam, e^{-itE(J)} psi0> = t sum (dE/dJ) Im(conj(lam) psi)."""
    norb, nelec, t = 4, (2, 2), 0.7
    dim = _dim(norb, nelec)
    lam, psi0 = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    j_aa = O.random_symmetric_uniform(norb, rng)
    j_ab = O.random_symmetric_uniform(norb, rng)

    def f(aa, ab):
        E = O.diag_coulomb_energies((aa, ab, aa), norb, nelec)
        return np.vdot(lam, np.exp(-1j * t * E) * psi0).real

    psi = np.exp(-1j * t * O.diag_coulomb_energies((j_aa, j_ab, j_aa), norb, nelec)) * psi0
    rsum, csum, M = O.diag_coulomb_phase_reductions(psi, lam, norb, nelec)
    occ_a = O.occupations(O.colex_strings(norb, nelec[0]), norb)
    occ_b = O.occupations(O.colex_strings(norb, nelec[1]), norb)
    g_aa = 0.5 * t * ((occ_a.T * rsum) @ occ_a + (occ_b.T * csum) @ occ_b)
    g_ab = t * occ_a.T @ M
    h = 1e-6
    for i, j in [(0, 0), (0, 1), (2, 3), (3, 1)]:
        E = np.zeros((norb, norb))
        E[i, j] = 1.0
        fd_aa = (f(j_aa + h * E, j_ab) - f(j_aa - h * E, j_ab)) / (2 * h)
        fd_ab = (f(j_aa, j_ab + h * E) - f(j_aa, j_ab - h * E)) / (2 * h)
        assert abs(fd_aa - g_aa[i, j]) < 1e-8
        assert abs(fd_ab - g_ab[i, j]) < 1e-8


def test_expm_adjoint_matches_finite_difference(rng):
    n = 5
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    K = A - A.conj().T
    K[0, 1] = 0.3j  # keep it anti-Hermitian with a near-degenerate pair
    K[1, 0] = 0.3j
    Q = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dK = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    dK = dK - dK.conj().T
    U = scipy.linalg.expm(K)
    h = 1e-6
    dU = (scipy.linalg.expm(K + h * dK) - scipy.linalg.expm(K - h * dK)) / (2 * h)
    lhs = np.trace(dU @ U.conj().T @ Q).real
    rhs = np.trace(dK @ _expm_adjoint(K, Q)).real
    assert abs(lhs - rhs) < 1e-8
    # degenerate spectrum (K = 0) reduces to R = Q
    np.testing.assert_allclose(_expm_adjoint(np.zeros((n, n), complex), Q), Q, atol=1e-14)


CASES = [
    # norb, nelec, n_reps, LUCJ pairs?, final rotation?
    (3, (2, 1), 1, False, False),
    (4, (2, 2), 2, True, True),
    (5, (3, 2), 2, False, True),
    (4, (1, 3), 3, True, False),
]


@pytest.mark.parametrize("norb, nelec, n_reps, lucj, final", CASES)
def test_adjoint_gradient_vs_oracle_finite_difference(norb, nelec, n_reps, lucj, final, rng):
    pairs = ([(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)]) if lucj else None
    layout = UCJLayout.make(norb, n_reps, pairs, final)
    n_params = O.ucj_n_params(norb, n_reps, pairs, final)
    assert layout.n_params == n_params
    theta = rng.uniform(-np.pi, np.pi, n_params)
    dim = _dim(norb, nelec)
    H = _random_hermitian(dim, rng)
    psi0 = O.random_state_vector(dim, rng)

    def energy(p):
        return O.ucj_energy(p, lambda v: H @ v, psi0, norb, nelec, n_reps, pairs, final)

    J, U, F = O.ucj_from_parameters(theta, norb, n_reps, pairs, final)
    psi = O.apply_ucj_spin_balanced(psi0, J, U, F, norb, nelec)
    ops = OracleAdjointOps(norb, nelec)
    grad, lam0 = ucj_adjoint_sweep(ops, layout, theta, psi, 2 * H @ psi)
    fd = O.fd_gradient(energy, theta, 1e-6)  # the reference's mechanism (h = 1e-6)
    scale = max(1.0, np.abs(fd).max())
    assert np.abs(grad - fd).max() < 1e-6 * scale
    # five-point differences reach ~1e-12 here: the adjoint holds to 1e-9
    fd5 = five_point_gradient(energy, theta)
    assert np.abs(grad - fd5).max() < 1e-9 * scale

    # gradient with respect to the reference state: W^dag (2 H psi)
    W = np.stack([O.apply_ucj_spin_balanced(e, J, U, F, norb, nelec) for e in np.eye(dim)], axis=1)
    np.testing.assert_allclose(lam0, W.conj().T @ (2 * H @ psi), atol=1e-10)


def test_adjoint_gradient_catches_planted_chain_rule_mistake(rng):
    """A transposed RDM in the chain rule moves the gradient off the FD value."""
    norb, nelec, n_reps = 4, (2, 2), 1
    layout = UCJLayout.make(norb, n_reps, None, True)
    theta = rng.uniform(-np.pi, np.pi, layout.n_params)
    dim = _dim(norb, nelec)
    H = _random_hermitian(dim, rng)
    psi0 = O.random_state_vector(dim, rng)
    J, U, F = O.ucj_from_parameters(theta, norb, n_reps, None, True)
    psi = O.apply_ucj_spin_balanced(psi0, J, U, F, norb, nelec)

    class Transposed(OracleAdjointOps):
        def transition_rdm(self, lam, psi):
            return super().transition_rdm(lam, psi).transpose(0, 2, 1)

    grad, _ = ucj_adjoint_sweep(Transposed(norb, nelec), layout, theta, psi, 2 * H @ psi)
    fd = O.fd_gradient(
        lambda p: O.ucj_energy(p, lambda v: H @ v, psi0, norb, nelec, n_reps, None, True),
        theta, 1e-6,
    )
    assert np.abs(grad - fd).max() > 1e-3


def test_layout_rejects_wrong_parameter_count():
    layout = UCJLayout.make(4, 2, None, False)
    with pytest.raises(ValueError):
        layout.unpack(np.zeros(layout.n_params + 1))
