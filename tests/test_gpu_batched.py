"""Batched spin-balanced UCJ applications (the serving path) against the oracle.

ffsim_apply_ucj_spin_balanced_batched applies a different operator to each
state of a batch of one shape.  Small states take the compound-matrix GEMM
path (psi <- Lambda^k(U_a) psi Lambda^k(U_b)^T with the merged rotations of
python/ffsim/variational/ucj_spin_balanced.py:657-696, batched diagonal
passes); larger ones the Givens path state by state.  Each result is compared
with the oracle's UCJ application of its own operator (complex128 1e-12,
complex64 2e-6), and with the single-state entry point.
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


@pytest.mark.parametrize("path", ["auto", "gemm", "sweep"])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize(
    "norb,nelec,n_states,final",
    [(4, (2, 1), 3, False), (6, (3, 3), 5, True), (8, (4, 4), 7, False), (8, (4, 3), 4, True),
     (10, (5, 5), 6, False), (10, (4, 5), 3, True), (12, (6, 6), 3, False), (13, (6, 5), 2, True)],
)
def test_batched_ucj_matches_oracle(norb, nelec, n_states, final, dtype, path, monkeypatch):
    """path: the library's choice, or FFSIM_AMD_BATCH_PATH forcing the GEMM path
    (taken up to 1024 strings per sector) or the Givens path."""
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
    if path == "sweep":
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
