"""GPU parity of UCJOpSpinUnbalanced / UCJOpSpinless against the oracle (1e-12).

Every test prints the case it runs (flushed) before touching the GPU.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(*args):
    print("CASE", *args, flush=True)
    sys.stdout.flush()


def _sym(n, rng):
    a = rng.uniform(-1, 1, (n, n))
    return a + a.T


@pytest.mark.parametrize("norb, nelec, final", [(4, (2, 1), True), (6, (3, 3), False),
                                                (8, (4, 3), True), (10, (5, 5), True)])
def test_ucj_spin_unbalanced(norb, nelec, final, rng):
    import ffsim_amd

    _case("ucj spin-unbalanced", norb, nelec, final)
    n_reps = 2
    n = ffsim_amd.UCJOpSpinUnbalanced.n_params(norb, n_reps, with_final_orbital_rotation=final)
    op = ffsim_amd.UCJOpSpinUnbalanced.from_parameters(
        rng.uniform(-np.pi, np.pi, n), norb=norb, n_reps=n_reps,
        with_final_orbital_rotation=final)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vec = O.random_state_vector(dim, rng)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    expected = O.apply_ucj_spin_unbalanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                           op.final_orbital_rotation, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)


@pytest.mark.parametrize("norb, nelec", [(6, 3), (10, 4), (8, (4, 2)), (12, (6, 6))])
def test_ucj_spinless(norb, nelec, rng):
    import ffsim_amd

    _case("ucj spinless", norb, nelec)
    n_reps = 3
    pairs = [(p, p + 1) for p in range(norb - 1)]
    n = ffsim_amd.UCJOpSpinless.n_params(norb, n_reps, interaction_pairs=pairs,
                                         with_final_orbital_rotation=True)
    op = ffsim_amd.UCJOpSpinless.from_parameters(
        rng.uniform(-np.pi, np.pi, n), norb=norb, n_reps=n_reps, interaction_pairs=pairs,
        with_final_orbital_rotation=True)
    sectors = (nelec, 0) if isinstance(nelec, int) else nelec
    dim = math.comb(norb, sectors[0]) * math.comb(norb, sectors[1])
    vec = O.random_state_vector(dim, rng)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    expected = O.apply_ucj_spinless(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                    op.final_orbital_rotation, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)


@pytest.mark.parametrize("norb, nelec", [(6, (3, 2)), (10, (5, 5))])
def test_ucj_angles_spin_balanced(norb, nelec, rng):
    import ffsim_amd
    from ffsim_amd.variational.givens import brickwork

    _case("ucj angles", norb, nelec)
    n_reps = 2
    nn_pairs = ([(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)])
    gv_pairs = list(brickwork(norb, norb))
    n = ffsim_amd.UCJAnglesOpSpinBalanced.n_params(norb, n_reps, nn_pairs, gv_pairs, True)
    op = ffsim_amd.UCJAnglesOpSpinBalanced.from_parameters(
        rng.uniform(-np.pi, np.pi, n), norb=norb, n_reps=n_reps,
        num_num_interaction_pairs=nn_pairs, givens_interaction_pairs=gv_pairs,
        with_final_givens_ansatz_op=True)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vec = O.random_state_vector(dim, rng)
    mats = np.stack([nn.to_diag_coulomb_mats() for nn in op.num_num_ansatz_ops])
    rots = np.stack([g.to_orbital_rotation() for g in op.givens_ansatz_ops])
    expected = O.apply_ucj_spin_balanced(vec, mats, rots,
                                         op.final_givens_ansatz_op.to_orbital_rotation(),
                                         norb, nelec)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
    # a Givens ansatz alone is the orbital rotation it describes
    g = op.givens_ansatz_ops[0]
    np.testing.assert_allclose(
        ffsim_amd.apply_unitary(vec, g, norb, nelec),
        O.apply_orbital_rotation(vec, g.to_orbital_rotation(), norb, nelec), atol=1e-12, rtol=0)


def _t2_real(norb, nocc, rng, scale=0.2):
    from ffsim_amd.random import random_t2_amplitudes

    return scale * random_t2_amplitudes(norb, nocc, seed=rng, dtype=float)


@pytest.mark.parametrize("norb, nelec, n_reps", [(8, (4, 4), None), (10, (5, 5), 3)])
def test_ucj_from_t_amplitudes(norb, nelec, n_reps, rng):
    """Operators built from t amplitudes (eigenvector rotations, rank-one
    diagonal Coulomb matrices) applied to the Hartree-Fock state and to a
    random state; the angle form gives the same state as the matrix form."""
    import ffsim_amd

    _case("ucj from t amplitudes", norb, nelec, n_reps)
    nocc = nelec[0]
    t2 = _t2_real(norb, nocc, rng)
    t1 = 0.1 * rng.standard_normal((nocc, norb - nocc))
    op = ffsim_amd.UCJOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=n_reps)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    hf = np.zeros(dim, dtype=complex)
    hf[0] = 1
    for vec in (hf, O.random_state_vector(dim, rng)):
        got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
        expected = O.apply_ucj_spin_balanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                             op.final_orbital_rotation, norb, nelec)
        np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
    angles = ffsim_amd.UCJAnglesOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=2)
    matrix = ffsim_amd.UCJOpSpinBalanced.from_t_amplitudes(t2, t1=t1, n_reps=2)
    np.testing.assert_allclose(ffsim_amd.apply_unitary(vec, angles, norb, nelec),
                               ffsim_amd.apply_unitary(vec, matrix, norb, nelec),
                               atol=1e-10, rtol=0)


def test_ucj_spin_unbalanced_from_t_amplitudes(rng):
    import ffsim_amd

    norb, nelec = 7, (4, 3)
    _case("ucj spin-unbalanced from t amplitudes", norb, nelec)
    t2aa = _t2_real(norb, nelec[0], rng)
    t2aa = t2aa - t2aa.transpose(1, 0, 2, 3)
    t2bb = _t2_real(norb, nelec[1], rng)
    t2bb = t2bb - t2bb.transpose(1, 0, 2, 3)
    t2ab = 0.2 * rng.standard_normal((nelec[0], nelec[1], norb - nelec[0], norb - nelec[1]))
    t1 = (0.1 * rng.standard_normal((nelec[0], norb - nelec[0])),
          0.1 * rng.standard_normal((nelec[1], norb - nelec[1])))
    op = ffsim_amd.UCJOpSpinUnbalanced.from_t_amplitudes((t2aa, t2ab, t2bb), t1=t1,
                                                         n_reps=(4, 2))
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vec = O.random_state_vector(dim, rng)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    expected = O.apply_ucj_spin_unbalanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                           op.final_orbital_rotation, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
