"""Serving scheduler on the CPU: admission, refill, failures, cancellation.

The scheduler runs over a fake backend defined here, so every failure path
(an error at launch, an error while running, an oversized request, a
cancelled request) is forced from the host and no GPU is involved.
"""

from __future__ import annotations

import numpy as np
import pytest

from ffsim_amd.serving import AdmissionError, LUCJRequest, LUCJServer, RequestCancelled


class FakeJob:
    def __init__(self, nbytes, rounds=2, fail_launch=False, fail_run=False):
        self.nbytes, self.rounds = nbytes, rounds
        self.fail_launch, self.fail_run = fail_launch, fail_run


class FakeBackend:
    """Each job runs for a fixed number of scheduler rounds on its stream."""

    def __init__(self, n_streams=2):
        self.n_streams = n_streams
        self.live_bytes = 0
        self.max_live_bytes = 0
        self.max_live_jobs = 0
        self.live = 0
        self.launched = []

    def footprint(self, req):
        return req.op.nbytes

    def launch(self, req, stream_index):
        job = req.op
        if job.fail_launch:
            raise RuntimeError("injected launch failure")
        self.live_bytes += job.nbytes
        self.live += 1
        self.max_live_bytes = max(self.max_live_bytes, self.live_bytes)
        self.max_live_jobs = max(self.max_live_jobs, self.live)
        self.launched.append(req)
        return {"job": job, "left": job.rounds, "stream": stream_index}

    def done(self, handle):
        if handle["job"].fail_run:
            raise RuntimeError("injected device failure")
        handle["left"] -= 1
        return handle["left"] <= 0

    def collect(self, req, handle):
        return ("result", id(req))

    def release(self, handle):
        self.live_bytes -= handle["job"].nbytes
        self.live -= 1


def _req(job):
    return LUCJRequest(op=job, norb=0, nelec=(0, 0))


def _server(budget=100, slots=2, n_streams=2, **kw):
    return LUCJServer(hbm_budget=budget, slots_per_stream=slots,
                      backend=FakeBackend(n_streams), poll_interval=0.0, **kw)


def test_budget_is_never_exceeded_and_everything_is_served():
    rng = np.random.default_rng(5)
    server = _server(budget=100)
    tickets = [server.submit(_req(FakeJob(int(rng.integers(5, 60)), int(rng.integers(1, 5)))))
               for _ in range(60)]
    while server.step():
        assert server.bytes_in_use <= 100
        assert server.backend.live_bytes <= 100
    assert all(t.result() == ("result", id(t.request)) for t in tickets)
    assert server.backend.max_live_bytes <= 100
    assert server.backend.max_live_jobs <= 4
    assert server.bytes_in_use == 0
    st = server.stats()
    assert st["served"] == 60 and st["failed"] == 0
    assert st["p50_ms"] <= st["p99_ms"]
    assert st["peak_hbm_bytes"] <= 100


def test_oversized_request_is_rejected_at_submission():
    server = _server(budget=100)
    with pytest.raises(AdmissionError):
        server.submit(_req(FakeJob(101)))
    t = server.submit(_req(FakeJob(100)))
    server.drain()
    assert t.result() is not None


def test_refill_keeps_slots_busy():
    server = _server(budget=10_000, slots=1, n_streams=2)
    for _ in range(10):
        server.submit(_req(FakeJob(1, rounds=3)))
    server.step()
    while True:
        n = server.step()
        if n == 0:
            break
        if n >= 2:  # while at least two are outstanding, both slots are in use
            assert len(server._inflight) == 2


def test_launch_failure_is_reported_and_memory_returned():
    server = _server(budget=100)
    good1 = server.submit(_req(FakeJob(30)))
    bad = server.submit(_req(FakeJob(30, fail_launch=True)))
    good2 = server.submit(_req(FakeJob(30)))
    server.drain()
    with pytest.raises(RuntimeError, match="injected launch failure"):
        bad.result()
    assert good1.result() and good2.result()
    assert server.bytes_in_use == 0
    assert server.stats()["failed"] == 1


def test_device_failure_is_reported_and_memory_returned():
    server = _server(budget=100)
    bad = server.submit(_req(FakeJob(50, fail_run=True)))
    good = server.submit(_req(FakeJob(50)))
    server.drain()
    with pytest.raises(RuntimeError, match="injected device failure"):
        bad.result()
    assert good.result()
    assert server.bytes_in_use == 0 and server.backend.live_bytes == 0


def test_cancel_only_before_admission():
    server = _server(budget=10, slots=1, n_streams=1)
    first = server.submit(_req(FakeJob(10, rounds=3)))
    queued = server.submit(_req(FakeJob(10)))
    server.step()  # admits the first only (budget)
    assert not server.cancel(first)
    assert server.cancel(queued)
    with pytest.raises(RequestCancelled):
        queued.result()
    server.drain()
    assert first.result()
    assert queued.request not in server.backend.launched


def test_large_request_is_not_starved():
    """Small requests may pass a large one only max_bypass times."""
    server = _server(budget=100, slots=4, n_streams=2, max_bypass=3)
    server.submit(_req(FakeJob(60, rounds=50)))  # occupies most of the budget
    server.step()
    big = server.submit(_req(FakeJob(80, rounds=1)))
    smalls = [server.submit(_req(FakeJob(10, rounds=50))) for _ in range(20)]
    server.drain()
    launched = server.backend.launched
    order = [r for r in launched if r is not launched[0]]
    assert order.index(big.request) <= 3
    assert big.result() and all(s.result() for s in smalls)


def test_background_thread_serves_concurrent_submissions():
    server = _server(budget=1000)
    server.start()
    try:
        tickets = [server.submit(_req(FakeJob(10, rounds=2))) for _ in range(50)]
        for t in tickets:
            assert t.result(timeout=30) == ("result", id(t.request))
    finally:
        server.stop()
    assert server.stats()["served"] == 50


class FakeBatchBackend(FakeBackend):
    """Batches requests of one key (the job's `key` attribute) into one launch."""

    def __init__(self, n_streams=2):
        super().__init__(n_streams)
        self.batches = []

    @staticmethod
    def batch_key(req):
        return req.op.key

    def launch_batch(self, reqs, stream_index):
        assert len({r.op.key for r in reqs}) == 1, "a batch mixes shapes"
        handles = [self.launch(r, stream_index) for r in reqs]
        self.batches.append([r.op.key for r in reqs])
        return {"handles": handles, "left": max(h["left"] for h in handles), "stream": stream_index,
                "job": reqs[0].op}

    def done(self, handle):
        if "handles" not in handle:
            return super().done(handle)
        for h in handle["handles"]:
            if h["job"].fail_run:
                raise RuntimeError("injected device failure")
        handle["left"] -= 1
        return handle["left"] <= 0

    def collect_batch(self, reqs, handle):
        return [("result", id(r)) for r in reqs]

    def release(self, handle):
        for h in handle["handles"]:
            super().release(h)


def test_batches_group_one_shape_and_respect_the_budget():
    rng = np.random.default_rng(9)
    backend = FakeBatchBackend(n_streams=2)
    server = LUCJServer(hbm_budget=400, slots_per_stream=2, backend=backend, poll_interval=0.0,
                        max_batch=5)
    jobs = []
    for _ in range(80):
        job = FakeJob(int(rng.integers(5, 40)), int(rng.integers(1, 4)))
        job.key = int(rng.integers(0, 3))
        jobs.append(job)
    tickets = [server.submit(_req(j)) for j in jobs]
    while server.step():
        assert server.bytes_in_use <= 400
        assert backend.live_bytes <= 400
    assert all(t.result() == ("result", id(t.request)) for t in tickets)
    assert all(len(b) <= 5 and len(set(b)) == 1 for b in backend.batches)
    assert max(len(b) for b in backend.batches) > 1  # requests were batched
    assert len(server.completed) == 80
