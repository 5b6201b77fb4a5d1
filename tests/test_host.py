"""Host-side checks that need no GPU: the C ABI library, the drop-in API
surface and the error behaviour of the reference's interface."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import ffsim_amd
from ffsim_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def _declared_symbols():
    text = (ROOT / "include/ffsim_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ffsim_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    declared = _declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    # and the Python binding types every one of them
    assert set(declared) == set(_lib.SIGNATURES)


def test_host_entry_points_without_gpu():
    assert _lib.lib().ffsim_amd_version() >= 1
    rots, diag = ffsim_amd.linalg.givens_decomposition(np.eye(3))
    assert rots == [] and np.allclose(diag, 1)


def test_public_api_names():
    for name in ["apply_orbital_rotation", "apply_diag_coulomb_evolution",
                 "apply_num_op_sum_evolution", "apply_unitary", "linear_operator",
                 "UCJOpSpinBalanced", "DoubleFactorizedHamiltonian", "DiagonalCoulombHamiltonian",
                 "hartree_fock_state", "dim", "dims"]:
        assert hasattr(ffsim_amd, name), name
    for name in ["contract_diag_coulomb", "diag_coulomb_linop", "contract_num_op_sum",
                 "num_op_sum_linop", "contract_one_body", "one_body_linop"]:
        assert hasattr(ffsim_amd.contract, name), name
    for name in ["givens_decomposition", "double_factorized_t2",
                 "double_factorized_t2_alpha_beta", "double_factorized", "modified_cholesky"]:
        assert hasattr(ffsim_amd.linalg, name), name
    for name in ["random_t2_amplitudes", "random_unitary", "random_ucj_op_spin_balanced"]:
        assert hasattr(ffsim_amd.random, name), name
    for cls in [ffsim_amd.UCJOpSpinBalanced, ffsim_amd.UCJOpSpinless,
                ffsim_amd.UCJOpSpinUnbalanced, ffsim_amd.UCJAnglesOpSpinBalanced]:
        assert hasattr(cls, "from_t_amplitudes"), cls
    assert hasattr(ffsim_amd.UCJOpSpinBalanced, "from_cisd_vec")
    assert hasattr(ffsim_amd.DoubleFactorizedHamiltonian, "from_molecular_hamiltonian")


def test_apply_unitary_type_error():
    # protocols/apply_unitary_protocol.py:81-88
    with pytest.raises(TypeError):
        ffsim_amd.apply_unitary(np.zeros(1), object(), 1, (1, 0))


def test_linear_operator_type_error():
    with pytest.raises(TypeError):
        ffsim_amd.linear_operator(object(), 2, (1, 1))


def test_ucj_spinless_not_implemented_returns_type_error():
    op = ffsim_amd.random.random_ucj_op_spin_balanced(3, n_reps=1, seed=1)
    with pytest.raises(TypeError):
        ffsim_amd.apply_unitary(np.zeros(3, dtype=complex), op, 3, 1)


def test_diag_coulomb_spinless_z_rep_not_implemented():
    with pytest.raises(NotImplementedError):
        ffsim_amd.apply_diag_coulomb_evolution(np.zeros(3, dtype=complex), np.eye(3), 1.0, 3, 1,
                                               z_representation=True)


def test_ucj_validation_errors():
    norb, n_reps = 3, 2
    good_J = np.zeros((n_reps, 2, norb, norb))
    good_U = np.stack([np.eye(norb, dtype=complex)] * n_reps)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(np.zeros((n_reps, 3, norb, norb)), good_U)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(good_J, good_U[0])
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(good_J, good_U[:1])
    bad = good_J.copy()
    bad[0, 0, 0, 1] = 1.0
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(bad, good_U)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(good_J, 2 * good_U)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced(good_J, good_U, final_orbital_rotation=2 * np.eye(norb))


def test_ucj_parameters_round_trip(rng):
    norb, n_reps = 4, 3
    pairs = ([(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)])
    # config "norb=12, n_reps=4, LUCJ pairs" has 4*(11+12+144) = 668 parameters
    assert ffsim_amd.UCJOpSpinBalanced.n_params(
        12, 4, interaction_pairs=([(p, p + 1) for p in range(11)], [(p, p) for p in range(12)])
    ) == 668
    for ip, final in ((None, False), (pairs, True)):
        n = ffsim_amd.UCJOpSpinBalanced.n_params(norb, n_reps, interaction_pairs=ip,
                                                 with_final_orbital_rotation=final)
        params = rng.uniform(-0.5, 0.5, n)
        op = ffsim_amd.UCJOpSpinBalanced.from_parameters(
            params, norb=norb, n_reps=n_reps, interaction_pairs=ip,
            with_final_orbital_rotation=final)
        np.testing.assert_allclose(op.to_parameters(interaction_pairs=ip), params, atol=1e-10)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced.from_parameters(np.zeros(3), norb=norb, n_reps=1)
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced.n_params(norb, 1, interaction_pairs=([(1, 0)], None))
    with pytest.raises(ValueError):
        ffsim_amd.UCJOpSpinBalanced.n_params(norb, 1, interaction_pairs=([(0, 1), (0, 1)], None))


def test_ucj_from_parameters_matches_oracle(rng):
    from oracle import oracle as O

    norb, n_reps = 4, 2
    n = ffsim_amd.UCJOpSpinBalanced.n_params(norb, n_reps, with_final_orbital_rotation=True)
    params = rng.uniform(-np.pi, np.pi, n)
    op = ffsim_amd.UCJOpSpinBalanced.from_parameters(params, norb=norb, n_reps=n_reps,
                                                     with_final_orbital_rotation=True)
    J, U, final = O.ucj_from_parameters(params, norb, n_reps, with_final_orbital_rotation=True)
    np.testing.assert_allclose(op.diag_coulomb_mats, J)
    np.testing.assert_allclose(op.orbital_rotations, U)
    np.testing.assert_allclose(op.final_orbital_rotation, final)


def test_random_generators_match_reference_draw_order():
    """Seeded generators follow python/ffsim/random/random.py draw order."""
    from oracle import oracle as O

    np.testing.assert_allclose(ffsim_amd.random.random_unitary(5, seed=11),
                               O.random_unitary(5, np.random.default_rng(11)))
    np.testing.assert_allclose(ffsim_amd.random.random_state_vector(7, seed=3),
                               O.random_state_vector(7, np.random.default_rng(3)))


def test_diagonal_coulomb_hamiltonian_diag_matches_oracle(rng):
    from oracle import oracle as O

    norb, nelec = 4, (2, 1)
    h = ffsim_amd.random.random_diagonal_coulomb_hamiltonian(norb, real=True, seed=rng)
    mats = (h.diag_coulomb_mats[0], h.diag_coulomb_mats[1], h.diag_coulomb_mats[0])
    expected = (O.diag_coulomb_energies(mats, norb, nelec)
                + O.num_op_energies(np.diag(h.one_body_tensor).real,
                                    np.diag(h.one_body_tensor).real, norb, nelec) + h.constant)
    np.testing.assert_allclose(h._diag_(norb, nelec), expected, atol=1e-12)


def test_gate_without_gpu_fails_loudly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="GPU"):
        ffsim_amd.apply_orbital_rotation(np.ones(4, dtype=complex) / 2, np.eye(2), 2, (1, 1))


def test_approx_eq_protocol(rng):
    """python/ffsim/protocols/approximate_equality_protocol.py:40-70: dispatch to
    either side's _approx_eq_, else ==."""
    op = ffsim_amd.random.random_ucj_op_spin_balanced(4, n_reps=2, seed=rng)
    same = ffsim_amd.UCJOpSpinBalanced(op.diag_coulomb_mats + 1e-10, op.orbital_rotations)
    assert ffsim_amd.approx_eq(op, same)
    assert not ffsim_amd.approx_eq(op, same, rtol=0, atol=1e-12)
    assert not ffsim_amd.approx_eq(op, "x") and not ffsim_amd.approx_eq("x", op)
    assert ffsim_amd.approx_eq(3, 3) and not ffsim_amd.approx_eq(3, 4)


def test_diag_and_trace_protocols(rng):
    """diagonal_protocol.py:40-50, trace_protocol.py:34-47."""
    from oracle import oracle as O

    norb, nelec = 4, (2, 1)
    h = ffsim_amd.random.random_diagonal_coulomb_hamiltonian(norb, real=True, seed=rng)
    d = ffsim_amd.diag(h, norb, nelec)
    np.testing.assert_allclose(d, h._diag_(norb, nelec))
    dense = (h.constant * np.eye(len(d)) + O.one_body_dense(h.one_body_tensor, norb, nelec)
             + np.diag(O.diag_coulomb_energies(
                 (h.diag_coulomb_mats[0], h.diag_coulomb_mats[1], h.diag_coulomb_mats[0]),
                 norb, nelec)))
    np.testing.assert_allclose(ffsim_amd.trace(h, norb, nelec), np.trace(dense).real, atol=1e-10)
    with pytest.raises(TypeError):
        ffsim_amd.diag(object(), norb, nelec)
    with pytest.raises(TypeError):
        ffsim_amd.trace(object(), norb, nelec)


def test_givens_layer_argument_errors_without_gpu():
    """The layer entry checks its rotations before touching the device: a
    non-adjacent pair is ffsim's AssertionError (gates/orbital_rotation.py:195),
    overlapping or out-of-range pairs are ValueErrors."""
    lib = _lib.lib()

    def call(i, j):
        n = len(i)
        c = np.full(n, 0.6)
        s = np.full(n, 0.8 + 0j)
        ii = np.asarray(i, dtype=np.int32)
        jj = np.asarray(j, dtype=np.int32)
        return _lib.check(lib.ffsim_apply_givens_layer(
            None, _lib.C128, 6, 3, 10, 1, n, c.ctypes.data_as(ctypes.c_void_p),
            s.ctypes.data_as(ctypes.c_void_p), ii.ctypes.data_as(ctypes.c_void_p),
            jj.ctypes.data_as(ctypes.c_void_p), None))

    with pytest.raises(AssertionError, match="adjacent"):
        call([0], [2])
    with pytest.raises(ValueError, match="disjoint"):
        call([0, 1], [1, 2])
    with pytest.raises(ValueError, match="out of range"):
        call([5], [6])


def test_batched_rotation_argument_errors_without_gpu():
    """The batched rotation checks its batch before touching the device."""
    import torch

    from ffsim_amd.gates.orbital_rotation import orbital_rotations_in_place

    norb, nelec = 4, (2, 2)
    states = [torch.zeros(36, dtype=torch.complex128) for _ in range(2)]
    u = np.eye(norb)
    with pytest.raises(ValueError, match="one rotation per state"):
        orbital_rotations_in_place(states, [u], norb, nelec)
    with pytest.raises(ValueError, match="same sectors"):
        orbital_rotations_in_place(states, [(u, None), (u, u)], norb, nelec)
    with pytest.raises(ValueError, match="one dtype"):
        orbital_rotations_in_place([states[0], states[1].to(torch.complex64)], [u, u], norb,
                                   nelec)
    with pytest.raises(ValueError, match="entries"):
        orbital_rotations_in_place([states[0], torch.zeros(35, dtype=torch.complex128)], [u, u],
                                   norb, nelec)
    with pytest.raises(ValueError, match="negative"):
        _lib.check(_lib.lib().ffsim_apply_orbital_rotation_batched(
            -1, None, None, _lib.C128, norb, 2, 2, None, None, None))


def test_batched_diag_argument_errors_without_gpu():
    """The batched diagonal evolution checks its batch before touching the device."""
    import torch

    from ffsim_amd.gates.diag_coulomb import diag_coulomb_evolutions_in_place

    norb, nelec = 4, (2, 2)
    j = np.eye(norb)
    with pytest.raises(ValueError, match="one dtype"):
        diag_coulomb_evolutions_in_place(
            [torch.zeros(36, dtype=torch.complex128), torch.zeros(36, dtype=torch.complex64)],
            j, 1.0, norb, nelec, False)
    with pytest.raises(ValueError, match="negative"):
        _lib.check(_lib.lib().ffsim_apply_diag_coulomb_evolution_batched(
            -1, None, _lib.C128, norb, 2, 2, None, None, None, 1.0, 0, None))


def test_batched_contraction_argument_errors_without_gpu():
    """The batched contraction checks its batch before touching the device."""
    import torch

    from ffsim_amd.contract.diag_coulomb import contract_diag_coulombs_in_place

    norb, nelec = 4, (2, 2)
    states = [torch.zeros(36, dtype=torch.complex128) for _ in range(2)]
    j = np.eye(norb)
    with pytest.raises(ValueError, match="one operator per state"):
        contract_diag_coulombs_in_place(states, [j], norb, nelec)
    with pytest.raises(ValueError, match="same interactions"):
        contract_diag_coulombs_in_place(states, [(j, None, j), (j, j, j)], norb, nelec)
    with pytest.raises(ValueError, match="negative"):
        _lib.check(_lib.lib().ffsim_contract_diag_coulomb_batched(
            -1, None, _lib.C128, norb, 2, 2, None, None, None, 0, 1.0, None))
