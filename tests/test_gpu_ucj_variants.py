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
