"""Batched spin-balanced UCJ applications (the serving path) against the oracle.

ffsim_apply_ucj_spin_balanced_batched applies a different operator to each
state of a batch of one shape.  Small states take the compound-matrix GEMM
path (psi <- Lambda^k(U_a) psi Lambda^k(U_b)^T with the merged rotations of
python/ffsim/variational/ucj_spin_balanced.py:657-696, batched diagonal
passes); larger ones the Givens path, with the tiles of all states in the
same launches where the beta sector fits a tile (else state by state).  Each
result is compared with the oracle's UCJ application of its own operator
(complex128 1e-12, complex64 2e-6), and with the single-state entry point,
which the Givens path matches bit for bit.
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


def _pairs(norb):
    return [(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)]


@pytest.mark.parametrize("path", ["auto", "gemm", "sweep", "states"])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize(
    "norb,nelec,n_states,final",
    [(4, (2, 1), 3, False), (6, (3, 3), 5, True), (8, (4, 4), 7, False), (8, (4, 3), 4, True),
     (10, (5, 5), 6, False), (10, (4, 5), 3, True), (12, (6, 6), 3, False), (13, (6, 5), 2, True),
     (9, (4, 0), 3, False)],
)
def test_batched_ucj_matches_oracle(norb, nelec, n_states, final, dtype, path, monkeypatch):
    """path: the library's choice, or FFSIM_AMD_BATCH_PATH forcing the GEMM path
    (taken up to 1024 strings per sector), the Givens path (all states in one
    launch per window pass where that applies) or the Givens path state by
    state."""
    import torch

    import ffsim_amd
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    rng = np.random.default_rng(100 * norb + n_states)
    ops = [ffsim_amd.random.random_ucj_op_spin_balanced(
        norb, n_reps=2, interaction_pairs=_pairs(norb), with_final_orbital_rotation=final,
        seed=rng) for _ in range(n_states)]
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vecs = [O.random_state_vector(dim, rng) for _ in range(n_states)]
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    if path != "auto":
        monkeypatch.setenv("FFSIM_AMD_BATCH_PATH", path)
    _case("batched ucj", norb, nelec, n_states, final, dtype, path)
    states = [torch.from_numpy(v).to("cuda").to(cdt) for v in vecs]
    single = [s.clone() for s in states]
    UCJOpSpinBalanced.apply_batched_in_place(ops, states, norb, nelec)
    for op, s in zip(ops, single):
        op.apply_in_place(s, norb, nelec)
    tol = 1e-12 if dtype == "c128" else 2e-6
    if path in ("sweep", "states"):
        # the Givens path runs each state exactly as the single-state entry does
        for got, one in zip(states, single):
            assert torch.equal(got, one)
    for op, v, got, one in zip(ops, vecs, states, single):
        expected = O.apply_ucj_spin_balanced(v, op.diag_coulomb_mats, op.orbital_rotations,
                                             op.final_orbital_rotation, norb, nelec)
        g = got.cpu().numpy().astype(np.complex128)
        np.testing.assert_allclose(g, expected, atol=tol, rtol=0)
        np.testing.assert_allclose(g, one.cpu().numpy().astype(np.complex128), atol=2 * tol,
                                   rtol=0)


def test_batched_ucj_rejects_mixed_operators():
    import torch

    import ffsim_amd
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    norb, nelec = 4, (2, 2)
    a = ffsim_amd.random.random_ucj_op_spin_balanced(norb, n_reps=1, seed=1)
    b = ffsim_amd.random.random_ucj_op_spin_balanced(norb, n_reps=2, seed=2)
    states = [torch.zeros(36, dtype=torch.complex128, device="cuda") for _ in range(2)]
    with pytest.raises(ValueError, match="n_reps"):
        UCJOpSpinBalanced.apply_batched_in_place([a, b], states, norb, nelec)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_batched_multi_state_mixed_structures(dtype, monkeypatch):
    """Over 256 states (two programs of the multi-state path), with a state
    whose merged rotation is the identity (every Givens angle vanishes, so its
    passes differ from the others') in the first program of one batch and in
    the second of another: the first batch runs state by state, the second
    queues its first program and runs the rest state by state.  Every result
    equals the single-state entry's bit for bit."""
    import torch

    import ffsim_amd
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    norb, nelec, n_states = 6, (3, 2), 300
    monkeypatch.setenv("FFSIM_AMD_BATCH_PATH", "sweep")
    rng = np.random.default_rng(7)
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    for odd in (3, 280):
        _case("batched ucj mixed structures", norb, nelec, n_states, odd, dtype)
        ops = [ffsim_amd.random.random_ucj_op_spin_balanced(
            norb, n_reps=2, interaction_pairs=_pairs(norb), seed=rng) for _ in range(n_states)]
        rots = ops[odd].orbital_rotations.copy()
        rots[1] = rots[0]  # merged rotation U_1^H U_0 = 1
        ops[odd] = UCJOpSpinBalanced(diag_coulomb_mats=ops[odd].diag_coulomb_mats,
                                     orbital_rotations=rots)
        vecs = [O.random_state_vector(dim, rng) for _ in range(n_states)]
        states = [torch.from_numpy(v).to("cuda").to(cdt) for v in vecs]
        single = [s.clone() for s in states]
        UCJOpSpinBalanced.apply_batched_in_place(ops, states, norb, nelec)
        for op, s in zip(ops, single):
            op.apply_in_place(s, norb, nelec)
        for got, one in zip(states, single):
            assert torch.equal(got, one)
        tol = 1e-12 if dtype == "c128" else 2e-6
        for k in (0, odd, n_states - 1):
            op = ops[k]
            expected = O.apply_ucj_spin_balanced(vecs[k], op.diag_coulomb_mats,
                                                 op.orbital_rotations, None, norb, nelec)
            np.testing.assert_allclose(states[k].cpu().numpy().astype(np.complex128), expected,
                                       atol=tol, rtol=0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("norb,nelec", [(5, (2, 2)), (8, (4, 4)), (12, (6, 5)), (13, (6, 6)),
                                        (8, (3, 0)), (7, (0, 2))])
@pytest.mark.parametrize("form", ["one", "alpha", "beta", "pair"])
def test_batched_orbital_rotation_bitwise(norb, nelec, form, dtype):
    """ffsim_apply_orbital_rotation_batched, each state with its own rotation:
    one matrix for both sectors, alpha only, beta only, or an (alpha, beta)
    pair.  norb 7, 8 and 12 share launches across the states; norb 5
    (windows narrower than a register sub-window) and norb 13 (beta sector
    transposed through one shared work buffer) run state by state.  Every result equals the single-state entry's
    bit for bit and the oracle's (complex128 1e-12, complex64 2e-6)."""
    import torch

    from ffsim_amd.gates.orbital_rotation import (
        orbital_rotation_in_place, orbital_rotations_in_place)

    rng = np.random.default_rng(norb * 10 + len(form))
    n_states = 5
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    _case("batched rotation", norb, nelec, form, dtype)

    def mat():
        return O.random_unitary(norb, rng)

    mats = []
    for _ in range(n_states):
        if form == "one":
            mats.append(mat())
        elif form == "alpha":
            mats.append((mat(), None))
        elif form == "beta":
            mats.append((None, mat()))
        else:
            mats.append((mat(), mat()))
    vecs = [O.random_state_vector(dim, rng) for _ in range(n_states)]
    states = [torch.from_numpy(v).to("cuda").to(cdt) for v in vecs]
    single = [s.clone() for s in states]
    orbital_rotations_in_place(states, mats, norb, nelec)
    for m, s in zip(mats, single):
        orbital_rotation_in_place(s, m, norb, nelec)
    for got, one in zip(states, single):
        assert torch.equal(got, one)
    tol = 1e-12 if dtype == "c128" else 2e-6
    for k in (0, n_states - 1):
        m = mats[k]
        if form == "alpha":
            m = (m[0], np.eye(norb))
        elif form == "beta":
            m = (np.eye(norb), m[1])
        expected = O.apply_orbital_rotation(vecs[k], m, norb, nelec)
        np.testing.assert_allclose(states[k].cpu().numpy().astype(np.complex128), expected,
                                   atol=tol, rtol=0)


def test_df_linop_batched_terms_match_term_by_term(monkeypatch):
    """The DF Hamiltonian's matvec on a small state (the terms' rotations in
    shared launches) equals the term-by-term loop bit for bit, and the oracle's
    DF matvec."""
    import torch

    import ffsim_amd
    from ffsim_amd.hamiltonians import double_factorized_hamiltonian as dfh

    norb, nelec = 8, (4, 3)
    rng = np.random.default_rng(5)
    ham = ffsim_amd.random.random_double_factorized_hamiltonian(norb, rank=4, seed=rng)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    v = O.random_state_vector(dim, rng)
    t = torch.from_numpy(v).to("cuda")
    _case("df linop batched terms", norb, nelec)
    lin = ffsim_amd.linear_operator(ham, norb, nelec)
    got = lin.matvec_device(t.clone())
    monkeypatch.setattr(dfh, "_batch_terms", lambda *a: False)  # the term-by-term loop
    lin_loop = ffsim_amd.linear_operator(ham, norb, nelec)
    ref = lin_loop.matvec_device(t.clone())
    assert torch.equal(got, ref)
    expected = O.df_hamiltonian_matvec(v, ham.one_body_tensor, ham.diag_coulomb_mats,
                                       ham.orbital_rotations, ham.constant, norb, nelec, False)
    np.testing.assert_allclose(got.cpu().numpy(), expected, atol=1e-10, rtol=0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("z_rep", [False, True])
@pytest.mark.parametrize("norb,nelec", [(6, (3, 2)), (12, (6, 6)), (14, (7, 7))])
def test_batched_diag_coulomb_evolution_bitwise(norb, nelec, z_rep, dtype):
    """ffsim_apply_diag_coulomb_evolution_batched: one operator on several
    states in one launch equals the single-state entry bit for bit, and the
    oracle (complex128 1e-12, complex64 2e-6)."""
    import torch

    from ffsim_amd.gates.diag_coulomb import (
        diag_coulomb_evolution_in_place, diag_coulomb_evolutions_in_place)

    rng = np.random.default_rng(norb + 100 * z_rep)
    _case("batched diag", norb, nelec, z_rep, dtype)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    j_aa = rng.standard_normal((norb, norb))
    j_aa = j_aa + j_aa.T
    j_ab = rng.standard_normal((norb, norb))
    j_ab = j_ab + j_ab.T
    mat = (j_aa, j_ab, j_aa)
    vecs = [O.random_state_vector(dim, rng) for _ in range(3)]
    states = [torch.from_numpy(v).to("cuda").to(cdt) for v in vecs]
    single = [s.clone() for s in states]
    diag_coulomb_evolutions_in_place(states, mat, 0.7, norb, nelec, z_rep)
    for s in single:
        diag_coulomb_evolution_in_place(s, mat, 0.7, norb, nelec, z_rep)
    for got, one in zip(states, single):
        assert torch.equal(got, one)
    expected = O.apply_diag_coulomb_evolution(vecs[1], mat, 0.7, norb, nelec,
                                              z_representation=z_rep)
    tol = 1e-12 if dtype == "c128" else 2e-6
    np.testing.assert_allclose(states[1].cpu().numpy().astype(np.complex128), expected,
                               atol=tol, rtol=0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("z_rep", [False, True])
def test_batched_diag_coulomb_contraction_bitwise(z_rep, dtype):
    """ffsim_contract_diag_coulomb_batched: each state contracted in place with
    its own operator equals the single-state contraction bit for bit, and the
    oracle (complex128 1e-12, complex64 relative 2e-6)."""
    import torch

    from ffsim_amd.contract.diag_coulomb import (
        contract_diag_coulomb_into, contract_diag_coulombs_in_place)

    norb, nelec = 9, (4, 5)
    rng = np.random.default_rng(31 + z_rep)
    _case("batched diag contraction", norb, nelec, z_rep, dtype)
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    mats = []
    for _ in range(4):
        m = rng.standard_normal((norb, norb))
        mats.append(m + m.T)
    vecs = [O.random_state_vector(dim, rng) for _ in range(4)]
    states = [torch.from_numpy(v).to("cuda").to(cdt) for v in vecs]
    single = [s.clone() for s in states]
    contract_diag_coulombs_in_place(states, mats, norb, nelec, z_rep, scale=0.5)
    for s, m in zip(single, mats):
        contract_diag_coulomb_into(s, s, m, norb, nelec, z_rep, scale=0.5)
    for got, one in zip(states, single):
        assert torch.equal(got, one)
    expected = 0.5 * O.contract_diag_coulomb(vecs[2], mats[2], norb, nelec,
                                             z_representation=z_rep)
    got = states[2].cpu().numpy().astype(np.complex128)
    if dtype == "c128":
        np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
    else:
        assert np.abs(got - expected).max() <= 2e-6 * np.abs(expected).max()


@pytest.mark.parametrize("order", ["clements", "triangle"])
def test_batched_multi_state_both_givens_orders(order, monkeypatch):
    """The multi-state launches under either Givens order (forced with
    FFSIM_AMD_GIVENS_ORDER): bitwise the single-state results, and the
    oracle's to 1e-12."""
    import torch

    import ffsim_amd
    from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

    norb, nelec, n_states = 10, (5, 4), 6
    monkeypatch.setenv("FFSIM_AMD_GIVENS_ORDER", order)
    monkeypatch.setenv("FFSIM_AMD_BATCH_PATH", "sweep")
    rng = np.random.default_rng(77)
    _case("batched ucj order", order, norb, nelec, n_states)
    ops = [ffsim_amd.random.random_ucj_op_spin_balanced(
        norb, n_reps=2, interaction_pairs=_pairs(norb), with_final_orbital_rotation=True,
        seed=rng) for _ in range(n_states)]
    dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    vecs = [O.random_state_vector(dim, rng) for _ in range(n_states)]
    states = [torch.from_numpy(v).to("cuda") for v in vecs]
    single = [s.clone() for s in states]
    UCJOpSpinBalanced.apply_batched_in_place(ops, states, norb, nelec)
    for op, s in zip(ops, single):
        op.apply_in_place(s, norb, nelec)
    for got, one in zip(states, single):
        assert torch.equal(got, one)
    op = ops[-1]
    expected = O.apply_ucj_spin_balanced(vecs[-1], op.diag_coulomb_mats, op.orbital_rotations,
                                         op.final_orbital_rotation, norb, nelec)
    np.testing.assert_allclose(states[-1].cpu().numpy(), expected, atol=1e-12, rtol=0)
