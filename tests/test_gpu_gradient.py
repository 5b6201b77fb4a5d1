"""GPU adjoint gradient of the UCJ energy against finite differences.

The reference's gradient is central finite differences
(python/ffsim/optimize/_util.py:60-110); the GPU adjoint sweep is checked
against (1) central differences of the GPU energy over every parameter,
(2) central differences of the CPU oracle's energy on a subset, and (3) the
same sweep run over the oracle's primitive steps.  The reductions it uses
(transition RDM, phase backward) are checked against the oracle directly.

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


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch


def _dim(norb, nelec):
    return math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])


def _lucj_pairs(norb):
    return [(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)]


@pytest.mark.parametrize("norb, nelec", [(3, (2, 1)), (6, (3, 2)), (8, (4, 4)), (9, (2, 6))])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_transition_rdm_and_phase_reductions(norb, nelec, dtype, rng):
    torch = _torch()
    from ffsim_amd.variational.gradient import DeviceAdjointOps

    _case("reductions", norb, nelec, dtype)
    tdt = torch.complex128 if dtype == "c128" else torch.complex64
    tol = 1e-12 if dtype == "c128" else 2e-5
    dim = _dim(norb, nelec)
    lam, psi = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    lam_d = torch.from_numpy(lam).to(tdt).cuda()
    psi_d = torch.from_numpy(psi).to(tdt).cuda()
    if dtype == "c64":  # compare against the oracle on the rounded inputs
        lam = lam_d.cpu().numpy().astype(complex)
        psi = psi_d.cpu().numpy().astype(complex)
    ops = DeviceAdjointOps(norb, nelec)
    rdm = ops.host_rdm(ops.transition_rdm(lam_d, psi_d))
    np.testing.assert_allclose(rdm, O.transition_rdm(lam, psi, norb, nelec), atol=tol, rtol=0)
    rsum, csum, M = ops.host_phase(ops.phase_reductions(psi_d, lam_d))
    e_rsum, e_csum, e_M = O.diag_coulomb_phase_reductions(psi, lam, norb, nelec)
    np.testing.assert_allclose(rsum, e_rsum, atol=tol, rtol=0)
    np.testing.assert_allclose(csum, e_csum, atol=tol, rtol=0)
    np.testing.assert_allclose(M, e_M, atol=tol, rtol=0)


@pytest.mark.parametrize("norb, nelec", [(14, (7, 6)), (20, (5, 1))])
def test_transition_rdm_large_sectors(norb, nelec, rng):
    """(14, (7, 6)): the beta sector's transposes (2 x 165 MB) exceed the call's
    scratch, so its strings are read along columns in place; (20, (5, 1)): the
    15504 alpha strings are reduced in several chunks."""
    torch = _torch()
    from ffsim_amd.variational.gradient import DeviceAdjointOps

    _case("trdm large", norb, nelec)
    dim = _dim(norb, nelec)
    lam, psi = O.random_state_vector(dim, rng), O.random_state_vector(dim, rng)
    ops = DeviceAdjointOps(norb, nelec)
    lam_d, psi_d = torch.from_numpy(lam).cuda(), torch.from_numpy(psi).cuda()
    rdm = ops.host_rdm(ops.transition_rdm(lam_d, psi_d))
    np.testing.assert_allclose(rdm, O.transition_rdm(lam, psi, norb, nelec), atol=1e-12, rtol=0)
    # the reduction order is fixed: a second call gives the same bits
    again = ops.host_rdm(ops.transition_rdm(lam_d, psi_d))
    np.testing.assert_array_equal(rdm, again)


def test_autograd_matches_oracle_finite_difference(rng):
    """loss.backward() through ucj_state and expectation, small case, all parameters."""
    torch = _torch()
    import ffsim_amd
    from ffsim_amd.variational.gradient import expectation, ucj_state

    norb, nelec, n_reps = 5, (3, 2), 2
    pairs = _lucj_pairs(norb)
    _case("autograd", norb, nelec, n_reps)
    n_params = O.ucj_n_params(norb, n_reps, pairs, True)
    theta = rng.uniform(-np.pi, np.pi, n_params)
    df = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=3, seed=rng)
    linop = ffsim_amd.linear_operator(df, norb, nelec)
    psi0 = O.random_state_vector(_dim(norb, nelec), rng)

    params = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    ref = torch.from_numpy(psi0).cuda().requires_grad_(True)
    psi = ucj_state(params, ref, norb=norb, nelec=nelec, n_reps=n_reps, interaction_pairs=pairs,
                    with_final_orbital_rotation=True)
    energy = expectation(psi, linop)
    energy.backward()

    def h(v):
        return O.df_hamiltonian_matvec(v, df.one_body_tensor, df.diag_coulomb_mats,
                                       df.orbital_rotations, df.constant, norb, nelec, False)

    e_ref = O.ucj_energy(theta, h, psi0, norb, nelec, n_reps, pairs, True)
    assert abs(energy.item() - e_ref) < 1e-10
    from oracle_adjoint import five_point_derivative, five_point_gradient

    fd = five_point_gradient(lambda p: O.ucj_energy(p, h, psi0, norb, nelec, n_reps, pairs, True),
                             theta)
    np.testing.assert_allclose(params.grad.numpy(), fd, atol=1e-9, rtol=0)
    # reference-state gradient: directional derivative along a random direction
    d = O.random_state_vector(psi0.size, rng)
    g_ref = ref.grad.cpu().numpy()
    J, U, F = O.ucj_from_parameters(theta, norb, n_reps, pairs, True)

    def e_of(v):
        w = O.apply_ucj_spin_balanced(v, J, U, F, norb, nelec)
        return np.vdot(w, h(w)).real

    fd_dir = five_point_derivative(lambda s: e_of(psi0 + s[0] * d), np.zeros(1), np.ones(1))
    assert abs(np.vdot(g_ref, d).real - fd_dir) < 1e-9


def test_config_norb12_gradient_668_params(rng):
    """norb=12 (6,6), LUCJ n_reps=4 (668 parameters, U(-pi, pi), seed 12), Hartree-Fock
    reference, synthetic H_DF with 8 factors (seed 7): adjoint vs five-point FD of the GPU
    energy on every parameter, vs five-point FD of the oracle's energy on a subset and
    along a random direction (all to 1e-9), and vs the sweep over the oracle's
    primitive steps (1e-10)."""
    torch = _torch()
    import ffsim_amd
    from ffsim_amd.variational.gradient import UCJLayout, ucj_adjoint_sweep, ucj_energy_and_gradient

    norb, nelec, n_reps = 12, (6, 6), 4
    pairs = _lucj_pairs(norb)
    n_params = O.ucj_n_params(norb, n_reps, pairs)
    assert n_params == 668
    theta = np.random.default_rng(12).uniform(-np.pi, np.pi, n_params)
    df = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=8, seed=7)
    linop = ffsim_amd.linear_operator(df, norb, nelec)
    psi0 = ffsim_amd.hartree_fock_state(norb, nelec)
    psi0_d = torch.from_numpy(psi0).cuda()
    _case("config norb=12 adjoint")
    energy, grad = ucj_energy_and_gradient(theta, linop, psi0_d, norb=norb, nelec=nelec,
                                           n_reps=n_reps, interaction_pairs=pairs)
    assert grad.shape == (668,)

    layout = UCJLayout.make(norb, n_reps, pairs)
    from ffsim_amd.variational.gradient import _apply_forward

    def gpu_energy(p):
        psi = psi0_d.clone()
        _apply_forward(layout, p, psi, nelec)
        return float(torch.vdot(psi, linop.matvec_device(psi)).real.item())

    from oracle_adjoint import five_point_derivative, five_point_gradient

    # Five-point central differences (h = 1e-3, truncation O(h^4)) reach about 1e-12 on
    # these energies (checked against the exact sweep over the oracle's steps on the
    # CPU), so the adjoint is held to 1e-9 absolute: a miss of 1e-7 fails.
    fd_tol = 1e-9
    _case("config norb=12 GPU five-point FD, 2672 energies")
    assert abs(gpu_energy(theta) - energy) < 1e-12
    fd_gpu = five_point_gradient(gpu_energy, theta)
    print(f"E={energy:.12f} max|adjoint-FD(GPU)|={np.abs(grad - fd_gpu).max():.2e}", flush=True)
    np.testing.assert_allclose(grad, fd_gpu, atol=fd_tol, rtol=0)

    def h(v):
        return O.df_hamiltonian_matvec(v, df.one_body_tensor, df.diag_coulomb_mats,
                                       df.orbital_rotations, df.constant, norb, nelec, False)

    def e_cpu_of(p):
        return O.ucj_energy(p, h, psi0, norb, nelec, n_reps, pairs)

    _case("config norb=12 oracle five-point FD on 6 parameters and one random direction")
    subset = [0, 70, 144, 155, 334, 667]
    fd_cpu = five_point_gradient(e_cpu_of, theta, indices=subset)
    print(f"max|adjoint-FD(oracle)| on {subset}: {np.abs(grad[subset] - fd_cpu).max():.2e}",
          flush=True)
    np.testing.assert_allclose(grad[subset], fd_cpu, atol=fd_tol, rtol=0)
    direction = np.random.default_rng(3).standard_normal(len(theta))
    direction /= np.linalg.norm(direction)
    d_cpu = five_point_derivative(e_cpu_of, theta, direction)
    print(f"directional: adjoint {grad @ direction:.15f} oracle FD {d_cpu:.15f}", flush=True)
    assert abs(grad @ direction - d_cpu) < fd_tol
    e_cpu = e_cpu_of(theta)
    assert abs(e_cpu - energy) < 1e-10

    _case("config norb=12 sweep over oracle steps")
    from oracle_adjoint import OracleAdjointOps

    J, U, F = O.ucj_from_parameters(theta, norb, n_reps, pairs)
    psi = O.apply_ucj_spin_balanced(psi0, J, U, F, norb, nelec)
    grad_cpu, _ = ucj_adjoint_sweep(OracleAdjointOps(norb, nelec), layout, theta, psi, 2 * h(psi))
    np.testing.assert_allclose(grad, grad_cpu, atol=1e-10, rtol=0)


def test_gradient_complex64_state_error(rng):
    """complex64 storage: the adjoint gradient's error against the complex128 one."""
    torch = _torch()
    import ffsim_amd
    from ffsim_amd.variational.gradient import ucj_energy_and_gradient

    norb, nelec, n_reps = 8, (4, 4), 2
    pairs = _lucj_pairs(norb)
    _case("gradient c64", norb, nelec)
    theta = rng.uniform(-np.pi, np.pi, O.ucj_n_params(norb, n_reps, pairs))
    df = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=4, seed=rng)
    linop = ffsim_amd.linear_operator(df, norb, nelec)
    psi0 = ffsim_amd.hartree_fock_state(norb, nelec)
    e128, g128 = ucj_energy_and_gradient(theta, linop, torch.from_numpy(psi0).cuda(), norb=norb,
                                         nelec=nelec, n_reps=n_reps, interaction_pairs=pairs)
    e64, g64 = ucj_energy_and_gradient(theta, linop, torch.from_numpy(psi0).to(torch.complex64).cuda(),
                                       norb=norb, nelec=nelec, n_reps=n_reps,
                                       interaction_pairs=pairs)
    err = np.abs(g64 - g128).max()
    print(f"c64 gradient: |dE| {abs(e64 - e128):.2e}, max |dgrad| {err:.2e}", flush=True)
    assert abs(e64 - e128) < 1e-5
    assert err < 1e-4 * max(1.0, np.abs(g128).max())
