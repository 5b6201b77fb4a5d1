"""Serving on the GPU: mixed shapes in flight together give the same results
as each request alone (within the stated tolerance: same-shape requests are
batched, and small states then take the compound-matrix GEMM path, whose
rounding differs from the Givens path's), and the HBM budget holds.

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


def _lucj_pairs(norb):
    return [(p, p + 1) for p in range(norb - 1)], [(p, p) for p in range(norb)]


def test_request_footprint_counts_state_and_work_buffer():
    from ffsim_amd.serving import request_footprint

    _case("footprint")
    for norb, nelec, dtype in [(8, (4, 4), "c128"), (12, (6, 5), "c64"), (14, (7, 7), "c128")]:
        elem = 16 if dtype == "c128" else 8
        n = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
        fp = request_footprint(norb, nelec, dtype)
        assert fp >= 2 * n * elem  # the state and a work buffer of its size
        assert fp < 2 * n * elem + 2 * (2 << 20) + 1


def _mixed_requests(rng, n, shapes):
    import ffsim_amd
    from ffsim_amd.serving import LUCJRequest

    reqs = []
    for k in range(n):
        norb, nelec, dtype = shapes[k % len(shapes)]
        op = ffsim_amd.random.random_ucj_op_spin_balanced(
            norb, n_reps=2, interaction_pairs=_lucj_pairs(norb), seed=rng
        )
        dim = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
        state = O.random_state_vector(dim, rng)
        reqs.append(LUCJRequest(op=op, norb=norb, nelec=nelec, state=state, dtype=dtype))
    return reqs


SHAPES = [(8, (4, 4), "c128"), (10, (5, 4), "c128"), (12, (6, 6), "c128"), (8, (4, 3), "c64"),
          (10, (5, 5), "c64"), (12, (6, 5), "c128")]


def test_results_identical_alone_and_in_mix(rng):
    import torch

    from ffsim_amd.serving import LUCJServer

    _case("serving mix", len(SHAPES))
    reqs = _mixed_requests(rng, 24, SHAPES)
    # alone: one request at a time on the default stream
    alone = []
    for r in reqs:
        cdt = torch.complex128 if r.dtype == "c128" else torch.complex64
        t = torch.from_numpy(r.state).to(cdt).cuda()
        r.op.apply_in_place(t, r.norb, r.nelec)
        alone.append(t.cpu().numpy())
    torch.cuda.synchronize()
    server = LUCJServer(hbm_budget=64 << 20, slots_per_stream=2, n_streams=4, max_batch=8)
    tickets = [server.submit(r) for r in reqs]
    server.drain(timeout=120)
    for r, t, a in zip(reqs, tickets, alone):
        got = t.result()
        assert got.dtype == a.dtype
        tol = 1e-12 if r.dtype == "c128" else 2e-6
        np.testing.assert_allclose(got, a, atol=tol, rtol=0)
    # and the c128 results match the oracle
    for r, t in zip(reqs[:6], tickets[:6]):
        if r.dtype != "c128":
            continue
        expected = O.apply_ucj_spin_balanced(r.state, r.op.diag_coulomb_mats,
                                             r.op.orbital_rotations, r.op.final_orbital_rotation,
                                             r.norb, r.nelec)
        np.testing.assert_allclose(t.result(), expected, atol=1e-12, rtol=0)
    st = server.stats()
    print("serving stats", st, flush=True)
    assert st["served"] == len(reqs) and st["failed"] == 0
    assert st["peak_hbm_bytes"] <= 64 << 20
    # the same requests in another mix (other batch compositions) agree as well
    server2 = LUCJServer(hbm_budget=64 << 20, slots_per_stream=1, n_streams=2, max_batch=3)
    tickets2 = [server2.submit(r) for r in reversed(reqs)]
    server2.drain(timeout=120)
    for r, t, t2 in zip(reqs, tickets, reversed(tickets2)):
        tol = 1e-12 if r.dtype == "c128" else 2e-6
        np.testing.assert_allclose(t2.result(), t.result(), atol=tol, rtol=0)


def test_batched_serving_is_bitwise_alone_on_the_givens_path(rng, monkeypatch):
    """With every batch on the Givens path, a request served in a batch gives
    exactly the bits it gives alone."""
    import torch

    from ffsim_amd.serving import LUCJServer

    monkeypatch.setenv("FFSIM_AMD_BATCH_PATH", "sweep")
    _case("serving bitwise sweep")
    reqs = _mixed_requests(rng, 12, SHAPES[:4])
    alone = []
    for r in reqs:
        cdt = torch.complex128 if r.dtype == "c128" else torch.complex64
        t = torch.from_numpy(r.state).to(cdt).cuda()
        r.op.apply_in_place(t, r.norb, r.nelec)
        alone.append(t.cpu().numpy())
    server = LUCJServer(hbm_budget=64 << 20, slots_per_stream=2, n_streams=2, max_batch=4)
    tickets = [server.submit(r) for r in reqs]
    server.drain(timeout=120)
    assert server.stats()["failed"] == 0
    for t, a in zip(tickets, alone):
        np.testing.assert_array_equal(t.result(), a)


def test_device_resident_requests_stay_on_the_device(rng):
    import torch

    from ffsim_amd.serving import LUCJRequest, LUCJServer

    _case("serving device-resident")
    reqs = _mixed_requests(rng, 6, SHAPES[:3])
    dev_reqs = []
    for r in reqs:
        cdt = torch.complex128 if r.dtype == "c128" else torch.complex64
        dev_reqs.append(LUCJRequest(op=r.op, norb=r.norb, nelec=r.nelec, dtype=r.dtype,
                                    state=torch.from_numpy(r.state).to(cdt).cuda(), to_host=False))
    server = LUCJServer(hbm_budget=64 << 20, slots_per_stream=2, n_streams=2)
    tickets = [server.submit(r) for r in dev_reqs]
    server.drain(timeout=120)
    for r, t in zip(reqs, tickets):
        got = t.result()
        assert isinstance(got, torch.Tensor) and got.device.type == "cuda"
        expected = O.apply_ucj_spin_balanced(r.state, r.op.diag_coulomb_mats,
                                             r.op.orbital_rotations, r.op.final_orbital_rotation,
                                             r.norb, r.nelec)
        np.testing.assert_allclose(got.cpu().numpy(), expected, atol=1e-12, rtol=0)


def test_tight_budget_serializes_large_requests(rng):
    from ffsim_amd.serving import AdmissionError, LUCJServer, request_footprint

    _case("serving tight budget")
    big = request_footprint(12, (6, 6), "c128")
    small = request_footprint(8, (4, 4), "c128")
    server = LUCJServer(hbm_budget=big + 2 * small, slots_per_stream=2, n_streams=2)
    reqs = _mixed_requests(rng, 8, [(12, (6, 6), "c128"), (8, (4, 4), "c128")])
    tickets = [server.submit(r) for r in reqs]
    server.drain(timeout=120)
    assert all(t.result() is not None for t in tickets)
    assert server.stats()["peak_hbm_bytes"] <= big + 2 * small
    with pytest.raises(AdmissionError):
        server.submit(_mixed_requests(rng, 1, [(14, (7, 7), "c128")])[0])
