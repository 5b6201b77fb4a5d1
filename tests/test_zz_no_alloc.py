"""The allocation-free diagonal Coulomb evolution against the regular entry point.

ffsim_apply_diag_coulomb_evolution_noalloc keeps its parameters, string
tables and column factors in a static device array instead of the library's
program ring and table cache.  Through the C ABI, two calls in a row (the
second on a smaller state) must give exactly the results of
ffsim_apply_diag_coulomb_evolution on the same inputs, the first call in a
fresh process must work before anything else of the library has run, and
repeated calls must not change the device's free memory.
"""

from __future__ import annotations

import ctypes
import math
import sys

import numpy as np
import pytest

from oracle import oracle as O


def _case(*args):
    print("CASE", *args, flush=True)
    sys.stdout.flush()


def _call(name, t, dtype_tag, norb, nelec, mats, time, z_rep):
    from ffsim_amd import _device, _lib

    aa, ab, bb = (None if m is None else np.ascontiguousarray(m, dtype=np.float64) for m in mats)
    _lib.call(name, _device.ptr(t), dtype_tag, norb, nelec[0], nelec[1], _device.host_ptr(aa),
              _device.host_ptr(ab), _device.host_ptr(bb), time, 1 if z_rep else 0,
              _device.stream_ptr())


def _prepared(J, z_rep):
    """aa / bb with the number representation's halved diagonal."""
    if z_rep:
        return J
    h = J.copy()
    h[np.diag_indices(len(J))] *= 0.5
    return h


CASES = [
    # (norb, nelec, dtype, z_rep): the second of each pair is the smaller state
    ((12, (6, 6), "c128", False), (8, (4, 3), "c128", False)),
    ((11, (5, 6), "c64", True), (6, (3, 3), "c64", True)),
    ((14, (7, 6), "c128", True), (9, (2, 7), "c64", False)),
]


@pytest.mark.gpu
@pytest.mark.parametrize("first,second", CASES)
def test_noalloc_matches_regular_entry_twice(first, second):
    import torch

    from ffsim_amd import _lib

    assert torch.cuda.is_available(), "GPU test"
    rng = np.random.default_rng(77)
    for norb, nelec, dtype, z_rep in (first, second):
        _case("noalloc", norb, nelec, dtype, z_rep)
        da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
        vec = O.random_state_vector(da * db, rng)
        J = O.random_symmetric_uniform(norb, rng)
        K = O.random_symmetric_uniform(norb, rng)
        X = rng.uniform(-1, 1, size=(norb, norb))
        mats = (_prepared(J, z_rep), X, _prepared(K, z_rep))
        cdt = torch.complex128 if dtype == "c128" else torch.complex64
        tag = _lib.C128 if dtype == "c128" else _lib.C64
        got = torch.from_numpy(vec).to("cuda").to(cdt)
        ref = got.clone()
        torch.cuda.synchronize()
        # allocation-free call first, so that in a fresh process it runs before
        # the library has set anything up
        _call("ffsim_apply_diag_coulomb_evolution_noalloc", got, tag, norb, nelec, mats, 0.7,
              z_rep)
        _call("ffsim_apply_diag_coulomb_evolution", ref, tag, norb, nelec, mats, 0.7, z_rep)
        torch.cuda.synchronize()
        assert torch.equal(got, ref), float((got - ref).abs().max())
        # and the oracle's closed form, to the usual bound
        e = O.diag_coulomb_energies((J, X, K), norb, nelec, z_representation=z_rep)
        expected = vec * np.exp(-0.7j * e)
        tol = 1e-12 if dtype == "c128" else 2e-6
        np.testing.assert_allclose(got.cpu().numpy().astype(np.complex128), expected, atol=tol,
                                   rtol=0)


@pytest.mark.gpu
def test_noalloc_leaves_device_memory_unchanged():
    import torch

    from ffsim_amd import _lib

    norb, nelec = 12, (6, 6)
    rng = np.random.default_rng(5)
    vec = O.random_state_vector(math.comb(12, 6) ** 2, rng)
    J = O.random_symmetric_uniform(norb, rng)
    mats = (_prepared(J, False), J, _prepared(J, False))
    t = torch.from_numpy(vec).to("cuda")
    _case("noalloc memory")
    _call("ffsim_apply_diag_coulomb_evolution_noalloc", t, _lib.C128, norb, nelec, mats, 0.3,
          False)  # warm-up: the runtime loads the kernels' code on first use
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for k in range(5):
        _call("ffsim_apply_diag_coulomb_evolution_noalloc", t, _lib.C128, norb, nelec, mats,
              0.1 * k, False)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0, (free0, free1)


def test_noalloc_rejects_states_beyond_its_scratch():
    """norb=22 (11,11): 705432 strings per sector need more than the 8 MiB of
    device scratch; the call fails with a ValueError before touching the GPU."""
    from ffsim_amd import _lib

    J = np.zeros((22, 22))
    lib = _lib.lib()
    status = lib.ffsim_apply_diag_coulomb_evolution_noalloc(
        None, _lib.C128, 22, 11, 11, J.ctypes.data_as(ctypes.c_void_p), None, None, 1.0, 0, None)
    with pytest.raises(ValueError, match="allocation-free"):
        _lib.check(status)
