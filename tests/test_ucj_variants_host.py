"""Spin-unbalanced and spinless UCJ: the oracle restatements pinned against
dense closed forms, and the host-side parameter layout of the drop-ins.

Pins: the oracle's ``apply_ucj_spin_unbalanced`` / ``apply_ucj_spinless`` are
checked against a product of compound-matrix (minors) orbital rotations and
eigenphase diagonal operators built from occupation numbers, which share no
code with the restated gate path.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as O
from ffsim_amd.variational.ucj_spin_unbalanced import UCJOpSpinUnbalanced
from ffsim_amd.variational.ucj_spinless import UCJOpSpinless


def _sym(n, rng):
    a = rng.uniform(-1, 1, (n, n))
    return a + a.T


def _closed_form_unbalanced(vec, mats, rots, final, norb, nelec):
    eye = np.eye(norb, dtype=complex)
    cur = (eye, eye)
    for m, r in zip(mats, rots):
        step = (r[0].conj().T @ cur[0], r[1].conj().T @ cur[1])
        vec = O.orbital_rotation_closed_form(vec, step[0], step[1], norb, nelec)
        E = O.diag_coulomb_energies((m[0], m[1], m[2]), norb, nelec)
        vec = np.exp(1j * E) * vec
        cur = (r[0], r[1])
    if final is not None:
        cur = (final[0] @ cur[0], final[1] @ cur[1])
    return O.orbital_rotation_closed_form(vec, cur[0], cur[1], norb, nelec)


@pytest.mark.parametrize("norb, nelec, final", [(3, (2, 1), True), (4, (2, 2), False),
                                                (4, (1, 3), True)])
def test_oracle_spin_unbalanced_vs_closed_form(norb, nelec, final, rng):
    n_reps = 2
    mats = np.stack([np.stack([_sym(norb, rng), rng.uniform(-1, 1, (norb, norb)), _sym(norb, rng)])
                     for _ in range(n_reps)])
    rots = np.stack([np.stack([O.random_unitary(norb, rng), O.random_unitary(norb, rng)])
                     for _ in range(n_reps)])
    fin = np.stack([O.random_unitary(norb, rng), O.random_unitary(norb, rng)]) if final else None
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vec = O.random_state_vector(dim, rng)
    got = O.apply_ucj_spin_unbalanced(vec, mats, rots, fin, norb, nelec)
    np.testing.assert_allclose(got, _closed_form_unbalanced(vec, mats, rots, fin, norb, nelec),
                               atol=1e-12)


@pytest.mark.parametrize("norb, nelec", [(4, 2), (5, 3), (4, (2, 1))])
def test_oracle_spinless_vs_closed_form(norb, nelec, rng):
    n_reps = 2
    mats = np.stack([_sym(norb, rng) for _ in range(n_reps)])
    rots = np.stack([O.random_unitary(norb, rng) for _ in range(n_reps)])
    fin = O.random_unitary(norb, rng)
    spinless = isinstance(nelec, int)
    sectors = (nelec, 0) if spinless else nelec
    dim = math.comb(norb, sectors[0]) * math.comb(norb, sectors[1])
    vec = O.random_state_vector(dim, rng)
    got = O.apply_ucj_spinless(vec, mats, rots, fin, norb, nelec)
    zero = np.zeros((norb, norb))
    cur = np.eye(norb, dtype=complex)
    v = vec
    for m, r in zip(mats, rots):
        step = r.conj().T @ cur
        v = O.orbital_rotation_closed_form(v, step, None if spinless else step, norb, sectors)
        E = O.diag_coulomb_energies((m, zero, zero if spinless else m), norb, sectors)
        v = np.exp(1j * E) * v
        cur = r
    cur = fin @ cur
    v = O.orbital_rotation_closed_form(v, cur, None if spinless else cur, norb, sectors)
    np.testing.assert_allclose(got, v, atol=1e-12)


def test_spin_unbalanced_parameter_round_trip(rng):
    norb, n_reps = 4, 2
    pairs = ([(0, 1), (1, 2)], [(0, 0), (1, 3), (3, 1)], [(2, 3)])
    for ip, final in ((None, True), (pairs, False)):
        n = UCJOpSpinUnbalanced.n_params(norb, n_reps, interaction_pairs=ip,
                                         with_final_orbital_rotation=final)
        expected = n_reps * ((10 if ip is None else 2) + (16 if ip is None else 3)
                             + (10 if ip is None else 1) + 2 * 16) + final * 32
        assert n == expected
        theta = rng.uniform(-0.5, 0.5, n)
        op = UCJOpSpinUnbalanced.from_parameters(theta, norb=norb, n_reps=n_reps,
                                                 interaction_pairs=ip,
                                                 with_final_orbital_rotation=final)
        np.testing.assert_allclose(op.to_parameters(interaction_pairs=ip), theta, atol=1e-10)
        with pytest.raises(ValueError):
            UCJOpSpinUnbalanced.from_parameters(theta[:-1], norb=norb, n_reps=n_reps,
                                                interaction_pairs=ip,
                                                with_final_orbital_rotation=final)
    with pytest.raises(ValueError):
        UCJOpSpinUnbalanced.n_params(norb, 1, interaction_pairs=([(1, 0)], None, None))
    with pytest.raises(ValueError):
        UCJOpSpinUnbalanced.n_params(norb, 1, interaction_pairs=(None, [(0, 1), (0, 1)], None))


def test_spinless_parameter_round_trip(rng):
    norb, n_reps = 5, 3
    for ip, final in ((None, True), ([(0, 1), (2, 2)], False)):
        n = UCJOpSpinless.n_params(norb, n_reps, interaction_pairs=ip,
                                   with_final_orbital_rotation=final)
        assert n == n_reps * ((15 if ip is None else 2) + 25) + final * 25
        theta = rng.uniform(-0.5, 0.5, n)
        op = UCJOpSpinless.from_parameters(theta, norb=norb, n_reps=n_reps, interaction_pairs=ip,
                                           with_final_orbital_rotation=final)
        np.testing.assert_allclose(op.to_parameters(interaction_pairs=ip), theta, atol=1e-10)


def test_validation_errors():
    with pytest.raises(ValueError):
        UCJOpSpinUnbalanced(diag_coulomb_mats=np.zeros((1, 2, 3, 3)),
                            orbital_rotations=np.zeros((1, 2, 3, 3)))
    with pytest.raises(ValueError):
        UCJOpSpinUnbalanced(diag_coulomb_mats=np.zeros((1, 3, 3, 3)),
                            orbital_rotations=np.stack([np.stack([np.eye(3), 2 * np.eye(3)])]))
    with pytest.raises(ValueError):
        UCJOpSpinless(diag_coulomb_mats=np.zeros((2, 3, 3)), orbital_rotations=np.zeros((1, 3, 3)))
    op = UCJOpSpinless(diag_coulomb_mats=np.zeros((1, 3, 3)),
                       orbital_rotations=np.eye(3)[None].astype(complex))
    assert op.norb == 3 and op.n_reps == 1
    assert UCJOpSpinUnbalanced(
        diag_coulomb_mats=np.zeros((1, 3, 3, 3)),
        orbital_rotations=np.stack([np.stack([np.eye(3), np.eye(3)])]).astype(complex),
    )._apply_unitary_(np.zeros(3), 3, 1, True) is NotImplemented


# ---- angle-parameterized operators (givens.py, num_num.py, ucj_angles_spin_balanced.py)

from ffsim_amd.variational.givens import GivensAnsatzOp, brickwork, brickwork_givens_rotations
from ffsim_amd.variational.num_num import NumNumAnsatzOpSpinBalanced
from ffsim_amd.variational.ucj_angles_spin_balanced import UCJAnglesOpSpinBalanced


@pytest.mark.parametrize("norb", [2, 3, 5, 8])
def test_givens_ansatz_round_trips_orbital_rotation(norb, rng):
    U = O.random_unitary(norb, rng)
    op = GivensAnsatzOp.from_orbital_rotation(U)
    assert all(abs(i - j) == 1 for i, j in op.interaction_pairs)
    assert len(op.interaction_pairs) == len(list(brickwork(norb, norb)))
    np.testing.assert_allclose(op.to_orbital_rotation(), U, atol=1e-12)
    n = GivensAnsatzOp.n_params(norb, op.interaction_pairs)
    assert n == 2 * len(op.interaction_pairs) + norb
    op2 = GivensAnsatzOp.from_parameters(op.to_parameters(), norb, op.interaction_pairs)
    np.testing.assert_allclose(op2.to_orbital_rotation(), U, atol=1e-12)


def test_givens_single_rotation_closed_form():
    """One rotation on (0, 1): columns (1, 0) map to (c x + s y, c y - conj(s) x) with
    s = sin(theta) e^{-i phi}, after the phase layer diag(e^{i alpha})."""
    theta, phi, alpha = 0.3, 1.1, np.array([0.2, -0.5, 0.7])
    op = GivensAnsatzOp(norb=3, interaction_pairs=[(0, 1)], thetas=np.array([theta]),
                        phis=np.array([phi]), phase_angles=alpha)
    d = np.exp(1j * alpha)
    c, s = np.cos(theta), np.sin(theta) * np.exp(-1j * phi)
    expected = np.diag(d).astype(complex)
    x, y = expected[:, 1].copy(), expected[:, 0].copy()
    expected[:, 1] = c * x + s * y
    expected[:, 0] = c * y - np.conj(s) * x
    np.testing.assert_allclose(op.to_orbital_rotation(), expected, atol=1e-15)
    assert abs(expected[0, 0] - c * d[0]) < 1e-15 and abs(expected[0, 1] - s * d[0]) < 1e-15


def test_brickwork_placement_preserves_the_product(rng):
    norb = 6
    pairs = [(0, 1), (2, 3), (1, 2), (4, 5), (3, 4), (1, 0), (2, 3)]
    thetas = list(rng.uniform(-1, 1, len(pairs)))
    phis = list(rng.uniform(-1, 1, len(pairs)))
    bp, bt, bph = brickwork_givens_rotations(pairs, thetas, phis, norb)
    a = GivensAnsatzOp(norb, pairs, np.array(thetas), np.array(phis), None).to_orbital_rotation()
    b = GivensAnsatzOp(norb, bp, np.array(bt), np.array(bph), None).to_orbital_rotation()
    np.testing.assert_allclose(a, b, atol=1e-14)


def test_num_num_ansatz_round_trips(rng):
    norb = 5
    J = np.stack([_sym(norb, rng), _sym(norb, rng)])
    op = NumNumAnsatzOpSpinBalanced.from_diag_coulomb_mats(J)
    np.testing.assert_allclose(op.to_diag_coulomb_mats(), J, atol=0)
    op2 = NumNumAnsatzOpSpinBalanced.from_parameters(op.to_parameters(), norb, op.interaction_pairs)
    np.testing.assert_allclose(op2.to_diag_coulomb_mats(), J, atol=0)
    with pytest.raises(ValueError):
        NumNumAnsatzOpSpinBalanced(norb, ([(1, 0)], []), (np.zeros(1), np.zeros(0)))


def test_ucj_angles_from_ucj_op_is_the_same_operator(rng):
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    norb, nelec, n_reps = 4, (2, 2), 2
    n = UCJOpSpinBalanced.n_params(norb, n_reps, with_final_orbital_rotation=True)
    ucj = UCJOpSpinBalanced.from_parameters(rng.uniform(-1, 1, n), norb=norb, n_reps=n_reps,
                                            with_final_orbital_rotation=True)
    ang = UCJAnglesOpSpinBalanced.from_ucj_op(ucj)
    assert ang.n_reps == n_reps
    vec = O.random_state_vector(36, rng)
    expected = O.apply_ucj_spin_balanced(vec, ucj.diag_coulomb_mats, ucj.orbital_rotations,
                                         ucj.final_orbital_rotation, norb, nelec)
    mats = np.stack([nn.to_diag_coulomb_mats() for nn in ang.num_num_ansatz_ops])
    rots = np.stack([g.to_orbital_rotation() for g in ang.givens_ansatz_ops])
    got = O.apply_ucj_spin_balanced(vec, mats, rots,
                                    ang.final_givens_ansatz_op.to_orbital_rotation(), norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12)


def test_ucj_angles_parameters(rng):
    norb, n_reps = 4, 2
    nn_pairs = ([(0, 1), (2, 3)], [(0, 0), (1, 1)])
    gv_pairs = list(brickwork(norb, 2))
    n = UCJAnglesOpSpinBalanced.n_params(norb, n_reps, nn_pairs, gv_pairs, True)
    assert n == n_reps * (4 + 2 * len(gv_pairs) + norb) + norb**2
    theta = rng.uniform(-1, 1, n)
    op = UCJAnglesOpSpinBalanced.from_parameters(theta, norb=norb, n_reps=n_reps,
                                                 num_num_interaction_pairs=nn_pairs,
                                                 givens_interaction_pairs=gv_pairs,
                                                 with_final_givens_ansatz_op=True)
    np.testing.assert_allclose(op.to_parameters(), theta, atol=0)
    with pytest.raises(ValueError):
        UCJAnglesOpSpinBalanced.from_parameters(theta[:-1], norb=norb, n_reps=n_reps,
                                                num_num_interaction_pairs=nn_pairs,
                                                givens_interaction_pairs=gv_pairs,
                                                with_final_givens_ansatz_op=True)
