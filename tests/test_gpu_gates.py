"""GPU parity of the gate drop-ins against the CPU oracle (complex128, 1e-12).

Every test prints the case it runs (flushed) before touching the GPU.
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
GOLDEN = Path(__file__).resolve().parent / "golden"


def _case(*args):
    print("CASE", *args, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch


def _dims(norb, nelec):
    return math.comb(norb, nelec[0]), math.comb(norb, nelec[1])


def _rand(norb, nelec, rng):
    da, db = _dims(norb, nelec)
    return O.random_state_vector(da * db, rng)


def test_library_loads_on_device():
    import ffsim_amd
    from ffsim_amd import _lib

    torch = _torch()
    _case("init")
    _lib.call("ffsim_amd_init", torch.cuda.current_device())
    assert ffsim_amd.dim(16, (8, 8)) == 12870**2


@pytest.mark.parametrize(
    "norb,nelec",
    [(2, (1, 1)), (3, (2, 1)), (3, (1, 2)), (4, (3, 2)), (5, (2, 3)), (6, (3, 3)), (8, (4, 4)),
     (8, (5, 5)), (8, (3, 5)), (10, (5, 5)), (10, (4, 6)), (12, (6, 6)), (13, (6, 5)),
     (14, (7, 7)), (14, (7, 6)), (4, (0, 4)), (4, (4, 0)), (5, (5, 2))],
)
def test_orbital_rotation_matches_oracle(norb, nelec, rng):
    import ffsim_amd

    _case("orbital_rotation", norb, nelec)
    vec = _rand(norb, nelec, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    for mat in (Ua, (Ua, Ub), (None, Ub), (Ua, None)):
        expected = O.apply_orbital_rotation(vec, mat, norb, nelec)
        got = ffsim_amd.apply_orbital_rotation(vec, mat, norb, nelec)
        np.testing.assert_allclose(got, expected, atol=TOL, rtol=0)


@pytest.mark.parametrize("order", ["clements", "triangle"])
@pytest.mark.parametrize("norb,nelec", [(8, (4, 3)), (13, (6, 5)), (14, (7, 7)), (15, (5, 6))])
def test_orbital_rotation_in_either_rotation_order(norb, nelec, order, rng, monkeypatch):
    """FFSIM_AMD_GIVENS_ORDER forces the rotation order of the fused passes
    (the library picks the one with fewer passes, then layer pairs): both give
    the oracle's result."""
    import ffsim_amd

    monkeypatch.setenv("FFSIM_AMD_GIVENS_ORDER", order)
    _case("orbital_rotation order", norb, nelec, order)
    vec = _rand(norb, nelec, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    expected = O.apply_orbital_rotation(vec, (Ua, Ub), norb, nelec)
    got = ffsim_amd.apply_orbital_rotation(vec, (Ua, Ub), norb, nelec)
    np.testing.assert_allclose(got, expected, atol=TOL, rtol=0)


@pytest.mark.parametrize("norb,nelec", [(0, (0, 0)), (1, (0, 0)), (1, (1, 0)), (1, (0, 1)),
                                        (1, (1, 1)), (2, (1, 1)), (2, (2, 2)), (3, (2, 1)),
                                        (3, (1, 2)), (4, (3, 2))])
def test_orbital_rotation_edge_cases(norb, nelec, rng):
    import ffsim_amd

    _case("orbital_rotation_edge", norb, nelec)
    da, db = _dims(norb, nelec)
    vec = O.random_state_vector(da * db, rng) if da * db else np.zeros(0, dtype=complex)
    U = O.random_unitary(norb, rng) if norb else np.zeros((0, 0), dtype=complex)
    got = ffsim_amd.apply_orbital_rotation(vec, U, norb, nelec)
    expected = O.orbital_rotation_closed_form(vec, U, U, norb, nelec) if da * db else vec
    np.testing.assert_allclose(got, expected, atol=TOL, rtol=0)


def test_orbital_rotation_special_case_fixture():
    """tests/python/gates/orbital_rotation_test.py:187-206 restated: the
    reference's regression unitary applied to the (5, 5) HF state."""
    import ffsim_amd

    _case("orbital_rotation_fixture")
    mat = np.load(GOLDEN / "orbital_rotation-0.npy")
    g = np.load(GOLDEN / "golden.npz")
    vec = ffsim_amd.hartree_fock_state(8, (5, 5))
    result = ffsim_amd.apply_orbital_rotation(vec, mat, 8, (5, 5))
    np.testing.assert_allclose(np.linalg.norm(result), 1, atol=1e-12)
    np.testing.assert_allclose(result, g["special_case_expected"], atol=1e-12, rtol=0)


def test_golden_config_norb8():
    import ffsim_amd

    _case("golden norb=8")
    g = np.load(GOLDEN / "golden.npz")
    norb, nelec, t = 8, (4, 4), float(g["time"])
    rotated = ffsim_amd.apply_orbital_rotation(g["vec"], g["U"], norb, nelec)
    np.testing.assert_allclose(rotated, g["rotated"], atol=TOL, rtol=0)
    dc = ffsim_amd.apply_diag_coulomb_evolution(rotated, g["J"], t, norb, nelec)
    np.testing.assert_allclose(dc, g["diag_coulomb"], atol=TOL, rtol=0)
    dcz = ffsim_amd.apply_diag_coulomb_evolution(rotated, g["J"], t, norb, nelec,
                                                 z_representation=True)
    np.testing.assert_allclose(dcz, g["diag_coulomb_z"], atol=TOL, rtol=0)
    numop = ffsim_amd.apply_num_op_sum_evolution(g["vec"], g["lam"], t, norb, nelec)
    np.testing.assert_allclose(numop, g["num_op"], atol=TOL, rtol=0)
    c = ffsim_amd.contract.contract_diag_coulomb(g["vec"], g["J"], norb, nelec)
    np.testing.assert_allclose(c, g["contract"], atol=TOL, rtol=0)
    cz = ffsim_amd.contract.contract_diag_coulomb(g["vec"], g["J"], norb, nelec,
                                                  z_representation=True)
    np.testing.assert_allclose(cz, g["contract_z"], atol=TOL, rtol=0)
    cn = ffsim_amd.contract.contract_num_op_sum(g["vec"], g["lam"], norb, nelec)
    np.testing.assert_allclose(cn, g["contract_num"], atol=TOL, rtol=0)


@pytest.mark.parametrize("z_rep", [False, True])
@pytest.mark.parametrize("norb,nelec", [(3, (2, 1)), (5, (2, 3)), (8, (4, 4)), (10, (5, 4))])
def test_diag_coulomb_evolution(norb, nelec, z_rep, rng):
    import ffsim_amd

    _case("diag_coulomb", norb, nelec, z_rep)
    vec = _rand(norb, nelec, rng)
    J = O.random_symmetric_uniform(norb, rng)
    t = 0.6
    np.testing.assert_allclose(
        ffsim_amd.apply_diag_coulomb_evolution(vec, J, t, norb, nelec, z_representation=z_rep),
        O.apply_diag_coulomb_evolution(vec, J, t, norb, nelec, z_representation=z_rep),
        atol=TOL, rtol=0,
    )
    mats = (O.random_symmetric_uniform(norb, rng), rng.standard_normal((norb, norb)),
            O.random_symmetric_uniform(norb, rng))
    for m in (mats, (None, mats[1], None), (mats[0], None, mats[2]), (None, None, None)):
        np.testing.assert_allclose(
            ffsim_amd.apply_diag_coulomb_evolution(vec, m, t, norb, nelec, z_representation=z_rep),
            O.apply_diag_coulomb_evolution(vec, m, t, norb, nelec, z_representation=z_rep),
            atol=TOL, rtol=0,
        )
    U = O.random_unitary(norb, rng)
    np.testing.assert_allclose(
        ffsim_amd.apply_diag_coulomb_evolution(vec, J, t, norb, nelec, orbital_rotation=U,
                                               z_representation=z_rep),
        O.apply_diag_coulomb_evolution(vec, J, t, norb, nelec, orbital_rotation=U,
                                       z_representation=z_rep),
        atol=TOL, rtol=0,
    )


def test_diag_coulomb_spinless(rng):
    import ffsim_amd

    norb, nelec = 6, 3
    _case("diag_coulomb spinless")
    vec = O.random_state_vector(math.comb(norb, nelec), rng)
    J = O.random_symmetric_uniform(norb, rng)
    np.testing.assert_allclose(
        ffsim_amd.apply_diag_coulomb_evolution(vec, J, 0.4, norb, nelec),
        O.apply_diag_coulomb_evolution(vec, J, 0.4, norb, nelec),
        atol=TOL, rtol=0,
    )
    with pytest.raises(NotImplementedError):
        ffsim_amd.apply_diag_coulomb_evolution(vec, J, 0.4, norb, nelec, z_representation=True)


@pytest.mark.parametrize("norb,nelec", [(4, (2, 2)), (7, (3, 4)), (9, (4, 4))])
def test_num_op_sum_evolution(norb, nelec, rng):
    import ffsim_amd

    _case("num_op", norb, nelec)
    vec = _rand(norb, nelec, rng)
    lam = rng.uniform(-1, 1, norb)
    lam_b = rng.uniform(-1, 1, norb)
    U = O.random_unitary(norb, rng)
    for coeffs, rot in ((lam, None), ((lam, lam_b), None), ((None, lam_b), None), (lam, U)):
        np.testing.assert_allclose(
            ffsim_amd.apply_num_op_sum_evolution(vec, coeffs, 0.9, norb, nelec, orbital_rotation=rot),
            O.apply_num_op_sum_evolution(vec, coeffs, 0.9, norb, nelec, orbital_rotation=rot),
            atol=TOL, rtol=0,
        )


@pytest.mark.parametrize("norb,nelec", [(4, (2, 2)), (8, (3, 5)), (11, (5, 5))])
def test_contractions_and_linops(norb, nelec, rng):
    import ffsim_amd

    _case("contractions", norb, nelec)
    vec = _rand(norb, nelec, rng)
    J = O.random_symmetric_uniform(norb, rng)
    U = O.random_unitary(norb, rng)
    for z in (False, True):
        np.testing.assert_allclose(
            ffsim_amd.contract.contract_diag_coulomb(vec, J, norb, nelec, z_representation=z),
            O.contract_diag_coulomb(vec, J, norb, nelec, z_representation=z),
            atol=TOL, rtol=0,
        )
        op = ffsim_amd.contract.diag_coulomb_linop(J, norb, nelec, orbital_rotation=U,
                                                   z_representation=z)
        np.testing.assert_allclose(
            op @ vec,
            O.diag_coulomb_matvec(vec, J, norb, nelec, orbital_rotation=U, z_representation=z),
            atol=TOL, rtol=0,
        )
    lam = rng.standard_normal(norb)
    op = ffsim_amd.contract.num_op_sum_linop(lam, norb, nelec, orbital_rotation=U)
    np.testing.assert_allclose(op @ vec, O.num_op_sum_matvec(vec, lam, norb, nelec,
                                                             orbital_rotation=U),
                               atol=TOL, rtol=0)


@pytest.mark.parametrize("norb,nelec,n_reps,final",
                         [(4, (2, 2), 1, False), (6, (3, 2), 2, True), (8, (4, 4), 3, True),
                          (12, (6, 6), 4, False), (14, (7, 7), 2, True), (13, (6, 5), 2, False),
                          (13, (4, 7), 1, True)])
def test_ucj_spin_balanced(norb, nelec, n_reps, final, rng):
    import ffsim_amd

    _case("ucj", norb, nelec, n_reps, final)
    op = ffsim_amd.random.random_ucj_op_spin_balanced(
        norb, n_reps=n_reps, with_final_orbital_rotation=final, seed=rng
    )
    vec = _rand(norb, nelec, rng)
    expected = O.apply_ucj_spin_balanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                         op.final_orbital_rotation, norb, nelec)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=TOL, rtol=0)


def test_df_hamiltonian_linop(rng):
    import ffsim_amd

    norb, nelec = 6, (3, 2)
    _case("df linop", norb, nelec)
    for z in (False, True):
        df = ffsim_amd.random.random_double_factorized_hamiltonian(norb, rank=4,
                                                                   z_representation=z, seed=rng)
        vec = _rand(norb, nelec, rng)
        op = ffsim_amd.linear_operator(df, norb, nelec)
        expected = O.df_hamiltonian_matvec(vec, df.one_body_tensor, df.diag_coulomb_mats,
                                           df.orbital_rotations, df.constant, norb, nelec, z)
        np.testing.assert_allclose(op @ vec, expected, atol=1e-10, rtol=0)
        # the other representation is the same operator
        other = df.to_number_representation() if z else df.to_z_representation()
        np.testing.assert_allclose(ffsim_amd.linear_operator(other, norb, nelec) @ vec,
                                   expected, atol=1e-10, rtol=0)


def test_diagonal_coulomb_hamiltonian_linop(rng):
    import ffsim_amd

    norb, nelec = 5, (2, 3)
    _case("dch linop")
    h = ffsim_amd.random.random_diagonal_coulomb_hamiltonian(norb, seed=rng)
    vec = _rand(norb, nelec, rng)
    op = ffsim_amd.linear_operator(h, norb, nelec)
    expected = O.diagonal_coulomb_hamiltonian_matvec(vec, h.one_body_tensor, h.diag_coulomb_mats,
                                                     h.constant, norb, nelec)
    np.testing.assert_allclose(op @ vec, expected, atol=1e-10, rtol=0)


def test_resident_tensor_in_place(rng):
    import ffsim_amd

    torch = _torch()
    norb, nelec = 10, (5, 5)
    _case("resident tensor")
    vec = _rand(norb, nelec, rng)
    U = O.random_unitary(norb, rng)
    t = torch.from_numpy(vec).cuda()
    out = ffsim_amd.apply_orbital_rotation(t, U, norb, nelec, copy=False)
    assert out.data_ptr() == t.data_ptr()
    t2 = torch.from_numpy(vec).cuda()
    out2 = ffsim_amd.apply_orbital_rotation(t2, U, norb, nelec, copy=True)
    assert out2.data_ptr() != t2.data_ptr()
    np.testing.assert_array_equal(t2.cpu().numpy(), vec)
    np.testing.assert_allclose(out.cpu().numpy(), O.apply_orbital_rotation(vec, U, norb, nelec),
                               atol=TOL, rtol=0)


@pytest.mark.parametrize("norb,nelec", [(8, (4, 4)), (12, (6, 6)), (13, (6, 5)), (14, (7, 6)),
                                        (14, (7, 7))])
def test_complex64_error_vs_fp64_oracle(norb, nelec, rng):
    """complex64 storage: the error against the complex128 oracle is stated
    per result and must stay at single-precision level (5e-5 absolute)."""
    import ffsim_amd

    torch = _torch()
    _case("c64", norb, nelec)
    vec = _rand(norb, nelec, rng)
    op = ffsim_amd.random.random_ucj_op_spin_balanced(norb, n_reps=2, seed=rng)
    expected = O.apply_ucj_spin_balanced(vec, op.diag_coulomb_mats, op.orbital_rotations,
                                         None, norb, nelec)
    t = torch.from_numpy(vec.astype(np.complex64)).cuda()
    got = ffsim_amd.apply_unitary(t, op, norb, nelec).cpu().numpy().astype(np.complex128)
    err = np.abs(got - expected).max()
    infidelity = 1 - abs(np.vdot(expected, got)) / np.linalg.norm(got)
    print(f"c64 norb={norb} |d|inf={err:.3e} 1-|<ref|psi>|={infidelity:.3e}", flush=True)
    assert err < 5e-5
    assert infidelity < 1e-6


@pytest.mark.parametrize("norb,nelec,order,n_steps,z_rep",
                         [(8, (4, 3), 0, 2, False), (8, (4, 3), 1, 2, False), (8, (4, 3), 2, 1, True),
                          (13, (6, 5), 1, 1, False), (13, (6, 5), 0, 1, True)])
def test_trotter_double_factorized(norb, nelec, order, n_steps, z_rep, rng):
    """norb=13 runs the beta sector transposed, where (number representation)
    the diagonal evolution is fused into the last pass of each rotation."""
    import ffsim_amd

    _case("trotter", norb, nelec, order, n_steps, z_rep)
    df = ffsim_amd.random.random_double_factorized_hamiltonian(norb, rank=5,
                                                               z_representation=z_rep, seed=rng)
    vec = _rand(norb, nelec, rng)
    expected = O.simulate_trotter_double_factorized(
        vec, df.one_body_tensor, df.diag_coulomb_mats, df.orbital_rotations, df.constant, 0.3,
        norb, nelec, n_steps=n_steps, order=order, z_representation=z_rep)
    got = ffsim_amd.simulate_trotter_double_factorized(vec, df, 0.3, norb=norb, nelec=nelec,
                                                       n_steps=n_steps, order=order)
    np.testing.assert_allclose(got, expected, atol=1e-11, rtol=0)


@pytest.mark.parametrize("norb,nelec,time,n_steps,order,atol",
                         [(3, (1, 1), 0.1, 20, 0, 1e-2), (4, (2, 1), 0.1, 10, 2, 1e-3),
                          (4, (2, 2), 0.1, 10, 1, 1e-3), (10, (5, 4), 0.1, 2, 1, None),
                          (13, (6, 5), 0.1, 2, 0, None)])
def test_trotter_diag_coulomb_split_op(norb, nelec, time, n_steps, order, atol, rng):
    """tests/python/trotter/diag_coulomb_split_op_test.py:24-83: against the
    exact propagator (expm_multiply through the GPU Hamiltonian linop) at the
    reference's tolerances for the small cases, and against the oracle at 1e-11
    everywhere; norb=13 runs the fused diagonal store."""
    import scipy.sparse.linalg

    import ffsim_amd

    _case("split-op trotter", norb, nelec, n_steps, order)
    h = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    h = h + h.conj().T
    aa, ab = (m @ m.T for m in (rng.standard_normal((norb, norb)) for _ in range(2)))
    ham = ffsim_amd.DiagonalCoulombHamiltonian(h, np.stack([aa, ab]),
                                               constant=float(rng.uniform(-10, 10)))
    vec = _rand(norb, nelec, rng)
    before = vec.copy()
    got = ffsim_amd.simulate_trotter_diag_coulomb_split_op(vec, ham, time, norb=norb,
                                                           nelec=nelec, n_steps=n_steps,
                                                           order=order)
    np.testing.assert_array_equal(vec, before)
    np.testing.assert_allclose(np.linalg.norm(got), 1.0, atol=1e-12)
    expected = O.simulate_trotter_diag_coulomb_split_op(vec, h, (aa, ab), ham.constant, time,
                                                        norb, nelec, n_steps=n_steps,
                                                        order=order)
    np.testing.assert_allclose(got, expected, atol=1e-11, rtol=0)
    if atol is not None:
        linop = ffsim_amd.linear_operator(ham, norb, nelec)
        exact = scipy.sparse.linalg.expm_multiply(
            -1j * time * linop, vec, traceA=-1j * time * np.sum(np.abs(ham.diag_coulomb_mats)))
        np.testing.assert_allclose(got, exact, atol=atol)


@pytest.mark.parametrize("norb,nelec", [(1, (1, 0)), (3, (2, 1)), (4, (0, 4)), (6, (3, 2)),
                                        (8, (4, 4)), (10, (5, 3)), (12, (6, 6))])
def test_contract_one_body(norb, nelec, rng):
    import ffsim_amd

    _case("one-body contraction", norb, nelec)
    vec = _rand(norb, nelec, rng)
    mat = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    expected = O.contract_one_body(vec, mat, norb, nelec)
    got = ffsim_amd.contract.contract_one_body(vec, mat, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
    op = ffsim_amd.contract.one_body_linop(mat, norb, nelec)
    np.testing.assert_allclose(op @ vec, expected, atol=1e-12, rtol=0)
    # rmatvec applies the conjugate transpose
    np.testing.assert_allclose(op.rmatvec(vec), O.contract_one_body(vec, mat.T.conj(), norb, nelec),
                               atol=1e-12, rtol=0)


def test_norb18_rotation_round_trip_and_trotter_norm():
    """Full-size norb=18 (9,9) state (48620^2, 37.8 GB complex128): an orbital
    rotation followed by its inverse returns the state, and a Trotter step
    preserves the norm (size-independent properties; the oracle cannot run here)."""
    torch = _torch()
    import ffsim_amd

    norb, nelec = 18, (9, 9)
    _case("norb=18 round trip")
    dim = math.comb(norb, 9) ** 2
    gen = torch.Generator(device="cuda")
    gen.manual_seed(18)
    psi = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=gen)
    psi /= torch.linalg.vector_norm(psi)
    ref = psi.clone()
    U = ffsim_amd.random.random_unitary(norb, seed=18)
    ffsim_amd.apply_orbital_rotation(psi, U, norb, nelec, copy=False)
    assert (psi - ref).abs().max().item() > 1e-6  # it did rotate
    ffsim_amd.apply_orbital_rotation(psi, U.T.conj(), norb, nelec, copy=False)
    assert (psi - ref).abs().max().item() < 1e-12
    del ref
    _case("norb=18 trotter step norm")
    ham = ffsim_amd.random.random_df_hamiltonian_normal(norb, n_factors=2, seed=18)
    ffsim_amd.simulate_trotter_double_factorized(psi, ham, 0.1, norb=norb, nelec=nelec, copy=False)
    assert abs(torch.linalg.vector_norm(psi).item() - 1.0) < 1e-12
    del psi
    torch.cuda.empty_cache()


@pytest.mark.parametrize("dtype,tol", [("c128", 1e-12), ("c64", 1e-8)])
def test_lucj_norb16_round_trip(dtype, tol):
    """Full-size LUCJ n_reps=4 at norb=16 (8,8) (12870^2, 2.65 GB in complex128):
    the operator followed by its inverse (reversed layers, negated diagonal
    Coulomb matrices) returns the state.  In complex128 this runs the fused
    diagonal store of every rotation at the benchmark's shape."""
    torch = _torch()
    import ffsim_amd

    norb, nelec = 16, (8, 8)
    _case("lucj norb=16 round trip", dtype)
    pairs = ([(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)])
    op = ffsim_amd.random.random_ucj_op_spin_balanced(norb, n_reps=4, interaction_pairs=pairs,
                                                      seed=16)
    inv = ffsim_amd.UCJOpSpinBalanced(
        diag_coulomb_mats=-op.diag_coulomb_mats[::-1].copy(),
        orbital_rotations=op.orbital_rotations[::-1].copy(),
    )
    dim = math.comb(norb, 8) ** 2
    gen = torch.Generator(device="cuda")
    gen.manual_seed(16)
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    psi = torch.randn(dim, dtype=torch.complex128, device="cuda", generator=gen)
    psi /= torch.linalg.vector_norm(psi)
    psi = psi.to(cdt)
    ref = psi.clone()
    op.apply_in_place(psi, norb, nelec)
    assert (psi - ref).abs().max().item() > 1e-6  # it did act
    inv.apply_in_place(psi, norb, nelec)
    err = (psi - ref).abs().max().item()
    print(f"lucj norb=16 {dtype} round trip |d|inf={err:.3e}", flush=True)
    assert err < tol
    del psi, ref
    torch.cuda.empty_cache()


def test_ucj_fused_diagonal_with_trivial_orbital_phases(rng):
    """Orbital rotations that are single real Givens rotations decompose with
    D exactly 1, so the fused diagonal store runs without orbital phases
    (norb=13: the beta sector is transposed)."""
    import ffsim_amd

    norb, nelec = 13, (6, 5)
    _case("ucj fused, trivial D", norb, nelec)
    rots = []
    for a, theta in ((2, 0.4), (9, -0.7)):
        U = np.eye(norb, dtype=complex)
        c, s = np.cos(theta), np.sin(theta)
        U[a, a], U[a, a + 1], U[a + 1, a], U[a + 1, a + 1] = c, s, -s, c
        _, d = ffsim_amd.linalg.givens_decomposition(U)
        assert np.all(d == 1)
        rots.append(U)
    ref = ffsim_amd.random.random_ucj_op_spin_balanced(norb, n_reps=2, seed=rng)
    op = ffsim_amd.UCJOpSpinBalanced(diag_coulomb_mats=ref.diag_coulomb_mats,
                                     orbital_rotations=np.stack(rots))
    vec = _rand(norb, nelec, rng)
    expected = O.apply_ucj_spin_balanced(vec, op.diag_coulomb_mats, op.orbital_rotations, None,
                                         norb, nelec)
    got = ffsim_amd.apply_unitary(vec, op, norb, nelec)
    np.testing.assert_allclose(got, expected, atol=TOL, rtol=0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("accumulate", [False, True])
def test_contract_one_body_hermitian_rotated_path(dtype, accumulate, rng):
    """A Hermitian M on a state that needs the transposition buffer (norb=13
    complex128, norb=15 complex64) runs as rotate / number operators / rotate
    back; against the oracle's link-table contraction, with scale and
    accumulation (complex128 1e-12; complex64 relative 2e-6)."""
    import torch

    from ffsim_amd.contract.one_body import _eigen_path, contract_one_body_into

    norb, nelec = (13, (6, 5)) if dtype == "c128" else (15, (4, 3))
    _case("one-body hermitian path", norb, nelec, dtype, accumulate)
    vec = _rand(norb, nelec, rng)
    a = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    mat = a + a.T.conj()
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    src = torch.from_numpy(vec).to("cuda").to(cdt)
    assert _eigen_path(np.ascontiguousarray(mat), src, norb, nelec)
    prev = _rand(norb, nelec, rng)
    out = torch.from_numpy(prev).to("cuda").to(cdt)
    contract_one_body_into(src, out, mat, norb, nelec, scale=0.5, accumulate=accumulate)
    expected = 0.5 * O.contract_one_body(vec, mat, norb, nelec) + (prev if accumulate else 0)
    got = out.cpu().numpy().astype(np.complex128)
    if dtype == "c128":
        np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
    else:
        assert np.abs(got - expected).max() <= 2e-6 * np.abs(expected).max()
    # a matrix that is not exactly Hermitian takes the window passes
    assert not _eigen_path(np.ascontiguousarray(mat + 1e-13j * np.eye(norb)), src, norb, nelec)
