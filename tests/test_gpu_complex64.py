"""complex64 states through every state entry point, against the complex128
oracle.  FP64 coefficients, FP32 storage and arithmetic: the error is stated
per result and must stay at single-precision level (states: 1e-6 absolute,
measured 1e-9 to 4e-8; operator products: 5e-6 relative to the largest
entry, measured up to 3.4e-7).

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


def _c64(vec):
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch.from_numpy(np.asarray(vec).astype(np.complex64)).cuda()


def _check_state(t, expected, what):
    import torch

    assert t.dtype == torch.complex64, what
    got = t.cpu().numpy().astype(np.complex128)
    err = np.abs(got - expected).max()
    print(f"c64 {what}: |d|inf={err:.2e}", flush=True)
    assert err < 1e-6, what


def _check_product(t, expected, what):
    import torch

    assert t.dtype == torch.complex64, what
    got = t.cpu().numpy().astype(np.complex128)
    err = np.abs(got - expected).max() / np.abs(expected).max()
    print(f"c64 {what}: rel |d|inf={err:.2e}", flush=True)
    assert err < 5e-6, what


def _vec(norb, nelec, rng):
    return O.random_state_vector(math.comb(norb, nelec[0]) * math.comb(norb, nelec[1]), rng)


@pytest.mark.parametrize("norb, nelec", [(6, (3, 2)), (10, (5, 4)), (13, (6, 5))])
def test_gates_complex64(norb, nelec, rng):
    import ffsim_amd

    _case("c64 gates", norb, nelec)
    vec = _vec(norb, nelec, rng)
    ua, ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    _check_state(ffsim_amd.apply_orbital_rotation(_c64(vec), (ua, ub), norb, nelec),
                 O.apply_orbital_rotation(vec, (ua, ub), norb, nelec), "orbital rotation")
    j = O.random_symmetric_uniform(norb, rng)
    for z in (False, True):
        _check_state(
            ffsim_amd.apply_diag_coulomb_evolution(_c64(vec), j, 0.8, norb, nelec,
                                                   orbital_rotation=ua, z_representation=z),
            O.apply_diag_coulomb_evolution(vec, j, 0.8, norb, nelec, orbital_rotation=ua,
                                           z_representation=z), f"diag coulomb z={z}")
    lam = rng.uniform(-1, 1, norb)
    _check_state(ffsim_amd.apply_num_op_sum_evolution(_c64(vec), lam, 0.7, norb, nelec,
                                                      orbital_rotation=ub),
                 O.apply_num_op_sum_evolution(vec, lam, 0.7, norb, nelec, orbital_rotation=ub),
                 "num-op sum")


@pytest.mark.parametrize("norb, nelec", [(6, (3, 2)), (10, (5, 4))])
def test_operators_complex64(norb, nelec, rng):
    import ffsim_amd

    _case("c64 operators", norb, nelec)
    vec = _vec(norb, nelec, rng)
    j = O.random_symmetric_uniform(norb, rng)
    lam = rng.uniform(-1, 1, norb)
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = h + h.conj().T
    _check_product(ffsim_amd.contract.diag_coulomb_linop(j, norb, nelec) @ _c64(vec),
                   O.contract_diag_coulomb(vec, j, norb, nelec), "diag-Coulomb linop")
    _check_product(ffsim_amd.contract.num_op_sum_linop(lam, norb, nelec) @ _c64(vec),
                   O.contract_num_op_sum(vec, lam, norb, nelec), "num-op linop")
    _check_product(ffsim_amd.contract.one_body_linop(h, norb, nelec) @ _c64(vec),
                   O.contract_one_body(vec, h, norb, nelec), "one-body linop")
    df = ffsim_amd.random.random_double_factorized_hamiltonian(norb, rank=3, seed=rng)
    _check_product(ffsim_amd.linear_operator(df, norb, nelec) @ _c64(vec),
                   O.df_hamiltonian_matvec(vec, df.one_body_tensor, df.diag_coulomb_mats,
                                           df.orbital_rotations, df.constant, norb, nelec,
                                           False), "DF Hamiltonian linop")


@pytest.mark.parametrize("norb, nelec", [(8, (4, 3)), (13, (6, 5))])
def test_sequences_complex64(norb, nelec, rng):
    import ffsim_amd

    _case("c64 sequences", norb, nelec)
    vec = _vec(norb, nelec, rng)
    df = ffsim_amd.random.random_double_factorized_hamiltonian(norb, rank=3, seed=rng)
    _check_state(
        ffsim_amd.simulate_trotter_double_factorized(_c64(vec), df, 0.3, norb=norb, nelec=nelec,
                                                     n_steps=2, order=1),
        O.simulate_trotter_double_factorized(vec, df.one_body_tensor, df.diag_coulomb_mats,
                                             df.orbital_rotations, df.constant, 0.3, norb,
                                             nelec, n_steps=2, order=1), "DF Trotter")
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = 0.5 * (h + h.conj().T)
    aa, ab = O.random_symmetric_uniform(norb, rng), O.random_symmetric_uniform(norb, rng)
    ham = ffsim_amd.DiagonalCoulombHamiltonian(h, np.stack([aa, ab]), constant=0.2)
    _check_state(
        ffsim_amd.simulate_trotter_diag_coulomb_split_op(_c64(vec), ham, 0.3, norb=norb,
                                                         nelec=nelec, n_steps=2),
        O.simulate_trotter_diag_coulomb_split_op(vec, h, (aa, ab), 0.2, 0.3, norb, nelec,
                                                 n_steps=2), "split-op Trotter")
    n = ffsim_amd.UCJOpSpinUnbalanced.n_params(norb, 2, with_final_orbital_rotation=True)
    op = ffsim_amd.UCJOpSpinUnbalanced.from_parameters(rng.uniform(-np.pi, np.pi, n), norb=norb,
                                                       n_reps=2,
                                                       with_final_orbital_rotation=True)
    _check_state(ffsim_amd.apply_unitary(_c64(vec), op, norb, nelec),
                 O.apply_ucj_spin_unbalanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                             op.final_orbital_rotation, norb, nelec),
                 "UCJ spin-unbalanced")
