"""GPU parity of ffsim's per-rotation native calls (ffsim_amd.native).

Each call is compared with the oracle's restatement of the same native call
(src/gates/orbital_rotation.rs:19-104, src/gates/phase_shift.rs:16-30) on the
same seeded inputs: complex128 to 1e-12, complex64 to 2e-6 (FP32 storage of
unit-scale amplitudes).  The last test drives ffsim's own per-rotation
schedule (python/ffsim/gates/orbital_rotation.py:117-154) through these calls
with the state resident on the GPU and compares it with the oracle's orbital
rotation.
"""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {"c128": 1e-12, "c64": 2e-6}


def _case(*args):
    print("CASE", *args, flush=True)
    sys.stdout.flush()


def _torch():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a ROCm GPU"
    return torch


def _upload(mat, dtype):
    torch = _torch()
    cdt = torch.complex128 if dtype == "c128" else torch.complex64
    return torch.from_numpy(np.ascontiguousarray(mat)).to("cuda").to(cdt)


def _pairs(norb, nocc, i, j):
    idx = O.zero_one_subspace_indices(norb, nocc, (i, j))
    half = len(idx) // 2
    return idx[:half], idx[half:]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize(
    "norb,nocc,i,j,n_cols",
    [(4, 2, 0, 1, 6), (6, 3, 2, 1, 20), (8, 4, 3, 4, 70), (10, 5, 8, 9, 300), (12, 6, 5, 6, 1),
     (12, 6, 0, 1, 1000)],
)
def test_givens_rotation_in_place(norb, nocc, i, j, n_cols, dtype, rng):
    from ffsim_amd import native

    _case("givens_rotation_in_place", norb, nocc, i, j, n_cols, dtype)
    dim = math.comb(norb, nocc)
    vec = O.random_state_vector(dim * n_cols, rng).reshape(dim, n_cols)
    c = float(rng.uniform(-1, 1))
    s = complex(np.sqrt(1 - c * c) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
    s1, s2 = _pairs(norb, nocc, i, j)
    expected = vec.copy()
    O._givens_in_place(expected, c, s, s1, s2)
    t = _upload(vec, dtype)
    native.apply_givens_rotation_in_place(t, c, s, s1, s2)
    got = t.cpu().numpy()
    np.testing.assert_allclose(got, expected, atol=TOL[dtype], rtol=0)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("norb,nocc,orb,n_cols", [(4, 2, 1, 6), (8, 4, 7, 70), (10, 3, 0, 120)])
def test_phase_shift_in_place(norb, nocc, orb, n_cols, dtype, rng):
    from ffsim_amd import native

    _case("phase_shift_in_place", norb, nocc, orb, n_cols, dtype)
    dim = math.comb(norb, nocc)
    vec = O.random_state_vector(dim * n_cols, rng).reshape(dim, n_cols)
    phase = complex(np.exp(1j * rng.uniform(0, 2 * np.pi)))
    idx = O.one_subspace_indices(norb, nocc, (orb,))
    expected = vec.copy()
    O._phase_shift_in_place(expected, phase, idx)
    t = _upload(vec, dtype)
    native.apply_phase_shift_in_place(t, phase, idx)
    np.testing.assert_allclose(t.cpu().numpy(), expected, atol=TOL[dtype], rtol=0)


def _random_layer(norb, rng, parity):
    rots = []
    for p in range(parity, norb - 1, 2):
        c = float(rng.uniform(-1, 1))
        s = complex(np.sqrt(1 - c * c) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
        # both orientations of an adjacent pair
        rots.append((c, s, p, p + 1) if rng.uniform() < 0.5 else (c, s, p + 1, p))
    return rots


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("norb,nelec", [(4, (2, 1)), (7, (3, 4)), (10, (5, 5)), (12, (6, 5))])
def test_givens_layer_both_sectors(norb, nelec, dtype, rng):
    """A layer on the alpha rows and on the beta columns equals the oracle's
    per-rotation calls (the beta one on the transposed state, as ffsim does)."""
    from ffsim_amd import native

    na, nb = nelec
    da, db = math.comb(norb, na), math.comb(norb, nb)
    vec = O.random_state_vector(da * db, rng).reshape(da, db)
    layer_a = _random_layer(norb, rng, 0)
    layer_b = _random_layer(norb, rng, 1)
    _case("givens_layer", norb, nelec, dtype, len(layer_a), len(layer_b))
    expected = vec.copy()
    for c, s, i, j in layer_a:
        O._givens_in_place(expected, c, s, *_pairs(norb, na, i, j))
    et = np.ascontiguousarray(expected.T)
    for c, s, i, j in layer_b:
        O._givens_in_place(et, c, s, *_pairs(norb, nb, i, j))
    expected = et.T
    t = _upload(vec, dtype)
    native.apply_givens_layer(t, layer_a, norb, na, strings_on_rows=True)
    native.apply_givens_layer(t, layer_b, norb, nb, strings_on_rows=False)
    np.testing.assert_allclose(t.cpu().numpy(), expected, atol=TOL[dtype], rtol=0)


@pytest.mark.parametrize("norb,nelec", [(6, (3, 2)), (9, (4, 5))])
def test_ffsim_schedule_through_native_calls(norb, nelec, rng):
    """ffsim's per-rotation schedule of apply_orbital_rotation, run call by call
    on the resident state (beta sector through the column form instead of the
    transpose copies), equals the oracle's orbital rotation."""
    from ffsim_amd import native

    na, nb = nelec
    da, db = math.comb(norb, na), math.comb(norb, nb)
    vec = O.random_state_vector(da * db, rng)
    ua, ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    _case("ffsim_schedule", norb, nelec)
    expected = O.apply_orbital_rotation(vec, (ua, ub), norb, nelec)
    t = _upload(vec.reshape(da, db), "c128")
    rots, phases = O.givens_decomposition(ua)
    for c, s, i, j in rots:
        s1, s2 = _pairs(norb, na, i, j)
        native.apply_givens_rotation_in_place(t, c, s.conjugate(), s1, s2)
    for i, ph in enumerate(phases):
        native.apply_phase_shift_in_place(t, ph, O.one_subspace_indices(norb, na, (i,)))
    rots, phases = O.givens_decomposition(ub)
    for c, s, i, j in rots:
        native.apply_givens_layer(t, [(c, s.conjugate(), i, j)], norb, nb, strings_on_rows=False)
    tt = t.T.contiguous()
    for i, ph in enumerate(phases):
        native.apply_phase_shift_in_place(tt, ph, O.one_subspace_indices(norb, nb, (i,)))
    got = tt.T.contiguous().cpu().numpy().reshape(-1)
    np.testing.assert_allclose(got, expected, atol=1e-12, rtol=0)
