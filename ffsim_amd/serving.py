"""Continuous-batching server for LUCJ applications of mixed shapes.

Requests (an operator, (norb, nelec), a reference state and a dtype) arrive
at any time.  The scheduler keeps a set of applications in flight on a few
HIP streams and refills it as soon as one finishes, admitting queued
requests in arrival order when both a stream slot and enough HBM are free:

* a request's footprint is its resident state plus the library's work
  buffer (``ffsim_orbital_rotation_work_elements``), rounded up to the
  allocator granule; the sum over requests in flight never exceeds the
  budget;
* a request larger than the whole budget is rejected at submission
  (:class:`AdmissionError`), never queued;
* a small request may pass a large one that does not fit yet, but only a
  bounded number of times, after which the queue head blocks admission until
  it fits (no starvation);
* the applications of different requests overlap on the device (one stream
  per slot group, default priority); each request's result is bitwise the
  same as when it runs alone, because every kernel is deterministic and a
  request's work stays on one stream.

* requests admitted in the same scheduling round that share a shape
  (norb, nelec, dtype, n_reps, final rotation or not) are launched together as
  one batch (``ffsim_apply_ucj_spin_balanced_batched``): small states take the
  compound-matrix GEMM path, one set of launches for the whole batch; larger
  ones run back to back on the batch's stream.

Latency is measured from arrival (submission) to completion seen by the
scheduler (event query); :meth:`LUCJServer.stats` reports p50/p99 and
throughput.  The device work itself is ``UCJOpSpinBalanced._apply_unitary_``
(python/ffsim/variational/ucj_spin_balanced.py:657-696) through the C ABI.
"""

from __future__ import annotations

import collections
import itertools
import math
import threading
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np

_GRANULE = 2 << 20  # torch's caching allocator rounds large blocks to 2 MiB


class AdmissionError(ValueError):
    """The request can never fit in the server's HBM budget."""


class RequestCancelled(RuntimeError):
    pass


@dataclass(eq=False)
class LUCJRequest:
    """One application of a spin-balanced UCJ operator to a state."""

    op: Any  # UCJOpSpinBalanced
    norb: int
    nelec: tuple
    state: Any = None  # numpy vector; None -> Hartree-Fock
    dtype: str = "c128"  # "c128" or "c64"
    to_host: bool = True  # copy the result back to a numpy vector


@dataclass(eq=False)
class Ticket:
    """Handle of a submitted request; ``result()`` blocks until it is served."""

    request: LUCJRequest
    id: int
    footprint: int
    t_arrival: float
    t_admit: float | None = None
    t_done: float | None = None
    bypassed: int = 0
    _value: Any = None
    _error: BaseException | None = None
    _event: threading.Event = field(default_factory=threading.Event)

    def done(self) -> bool:
        return self._event.is_set()

    def result(self, timeout: float | None = None):
        if not self._event.wait(timeout):
            raise TimeoutError(f"request {self.id} not served within {timeout} s")
        if self._error is not None:
            raise self._error
        return self._value

    @property
    def latency(self) -> float | None:
        return None if self.t_done is None else self.t_done - self.t_arrival

    def _finish(self, value=None, error=None, now=None):
        self._value, self._error = value, error
        self.t_done = time.perf_counter() if now is None else now
        self._event.set()


def request_footprint(norb: int, nelec, dtype: str) -> int:
    """HBM bytes a request holds while in flight: its state and a work buffer of
    the state's size (the batched path's GEMMs run out of place), each rounded up
    to the allocator granule."""
    elem = 16 if dtype == "c128" else 8
    n = math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])
    return 2 * (-(-(n * elem) // _GRANULE) * _GRANULE)


def batch_key(req: "LUCJRequest"):
    """Requests with equal keys can share one batched launch."""
    return (req.norb, tuple(req.nelec), req.dtype, req.op.n_reps,
            req.op.final_orbital_rotation is not None)


@dataclass
class _Launch:
    """Device work of one launch: its states (results), their work buffers, the
    event recorded behind it, its stream and the pinned copies of host results."""

    states: list
    works: list
    event: Any
    stream: Any
    staged: list


class DeviceBackend:
    """Runs requests on the GPU: one torch stream per slot group."""

    def __init__(self, n_streams: int = 4):
        import torch

        from ffsim_amd import _device

        self.torch = torch
        self._device = _device
        self.device = _device.require_device()
        # default priority only; at most the number of hardware queues
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(n_streams)]

    @property
    def n_streams(self) -> int:
        return len(self.streams)

    def footprint(self, req: LUCJRequest) -> int:
        return request_footprint(req.norb, tuple(req.nelec), req.dtype)

    def _device_state(self, req: LUCJRequest):
        torch = self.torch
        from ffsim_amd.states import hartree_fock_state

        cdt = torch.complex128 if req.dtype == "c128" else torch.complex64
        if isinstance(req.state, torch.Tensor) and req.state.device.type == "cuda":
            return req.state.reshape(-1).to(cdt, copy=True)
        host = req.state if req.state is not None else hartree_fock_state(req.norb, req.nelec)
        host = torch.from_numpy(np.ascontiguousarray(host, dtype=np.complex128)).to(cdt)
        return self._device.host_to_device(host.numpy(), self.device)

    batch_key = staticmethod(batch_key)

    def launch_batch(self, reqs: list, stream_index: int):
        """One batched application for requests of one shape (see batch_key)."""
        torch = self.torch
        from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

        r0 = reqs[0]
        stream = self.streams[stream_index]
        with torch.cuda.stream(stream):
            states = [self._device_state(r) for r in reqs]
            works = [torch.empty_like(t) for t in states]
            UCJOpSpinBalanced.apply_batched_in_place([r.op for r in reqs], states, r0.norb,
                                                     tuple(r0.nelec), works)
            staged = [self._device.start_device_to_host(t) if r.to_host else None
                      for r, t in zip(reqs, states)]
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Launch(states, works, ev, stream, staged)

    def collect_batch(self, reqs: list, handle: _Launch) -> list:
        handle.event.synchronize()
        out = []
        for t, st in zip(handle.states, handle.staged):
            if st is not None:
                out.append(self._device.finish_device_to_host(st))
            else:
                t.record_stream(self.torch.cuda.default_stream(self.device))
                out.append(t)
        return out

    def launch(self, req: LUCJRequest, stream_index: int):
        """One request alone (the unbatched form of launch_batch)."""
        torch = self.torch

        stream = self.streams[stream_index]
        with torch.cuda.stream(stream):
            t = self._device_state(req)
            req.op.apply_in_place(t, req.norb, tuple(req.nelec))
            # the result's copy to pinned host memory is queued behind the kernels,
            # so a finished request needs only a host-side copy to be collected
            staged = self._device.start_device_to_host(t) if req.to_host else None
            ev = torch.cuda.Event()
            ev.record(stream)
        return _Launch([t], [], ev, stream, [staged])

    def done(self, handle: _Launch) -> bool:
        return handle.event.query()

    def collect(self, req: LUCJRequest, handle: _Launch):
        """The result: a numpy vector, or (``to_host=False``) the device tensor.

        A device result was allocated on the server's side stream and is complete
        when returned.  It is marked as used by the device's default stream, so
        the caching allocator does not hand its memory to another request while
        work queued there still reads it; a caller that reads it on yet another
        stream calls ``tensor.record_stream(that_stream)`` first.
        """
        return self.collect_batch([req], handle)[0]

    def release(self, handle) -> None:
        pass


class LUCJServer:
    """Scheduler: admission by HBM footprint, continuous refill of in-flight slots.

    Each slot holds one launch: a single request, or, with a backend that
    batches (``launch_batch``), up to ``max_batch`` requests of one shape
    admitted in the same scheduling round."""

    def __init__(self, *, hbm_budget: int, slots_per_stream: int = 2, backend=None,
                 n_streams: int = 4, max_bypass: int = 8, poll_interval: float = 2e-5,
                 max_batch: int = 32):
        self.backend = backend if backend is not None else DeviceBackend(n_streams)
        self.hbm_budget = int(hbm_budget)
        self.batching = hasattr(self.backend, "launch_batch")
        self.max_batch = max(1, int(max_batch)) if self.batching else 1
        self.max_launches = slots_per_stream * self.backend.n_streams
        self.max_inflight = self.max_launches * self.max_batch
        self.max_bypass = max_bypass
        self.poll_interval = poll_interval
        self._lock = threading.Lock()
        self._ids = itertools.count()
        self._queue: collections.deque[Ticket] = collections.deque()
        # (tickets of one launch, backend handle, stream index)
        self._inflight: list[tuple[list[Ticket], Any, int]] = []
        self._stream_load = [0] * self.backend.n_streams
        self.bytes_in_use = 0
        self.peak_bytes = 0
        self.completed: list[Ticket] = []
        self._thread: threading.Thread | None = None
        self._stop = threading.Event()

    # -- client side -------------------------------------------------------
    def submit(self, req: LUCJRequest, now: float | None = None) -> Ticket:
        fp = self.backend.footprint(req)
        if fp > self.hbm_budget:
            raise AdmissionError(
                f"request needs {fp} bytes of HBM but the server budget is {self.hbm_budget}"
            )
        t = Ticket(req, next(self._ids), fp, time.perf_counter() if now is None else now)
        with self._lock:
            self._queue.append(t)
        return t

    def cancel(self, ticket: Ticket) -> bool:
        """Cancel a request that has not been admitted yet."""
        with self._lock:
            try:
                self._queue.remove(ticket)
            except ValueError:
                return False
        ticket._finish(error=RequestCancelled(f"request {ticket.id} cancelled"))
        return True

    # -- scheduler ---------------------------------------------------------
    def _in_flight_requests(self) -> int:
        return sum(len(g) for g, _, _ in self._inflight)

    def _admit(self) -> list[Ticket]:
        admitted = []
        with self._lock:
            free_slots = self.max_inflight - self._in_flight_requests()
            kept = collections.deque()
            waiting: list[Ticket] = []  # earlier requests that do not fit yet
            while self._queue and free_slots > 0:
                t = self._queue.popleft()
                if any(w.bypassed >= self.max_bypass for w in waiting):
                    kept.append(t)  # an old request blocks admission until it fits
                    continue
                if self.bytes_in_use + t.footprint <= self.hbm_budget:
                    self.bytes_in_use += t.footprint
                    self.peak_bytes = max(self.peak_bytes, self.bytes_in_use)
                    admitted.append(t)
                    free_slots -= 1
                    for w in waiting:
                        w.bypassed += 1
                else:
                    waiting.append(t)
                    kept.append(t)
            kept.extend(self._queue)
            self._queue = kept
        return admitted

    def _groups(self, tickets: list[Ticket]) -> list[list[Ticket]]:
        """Launch groups: one per request, or same-shape batches of at most max_batch."""
        if not self.batching:
            return [[t] for t in tickets]
        by_key: dict = {}
        for t in tickets:
            by_key.setdefault(self.backend.batch_key(t.request), []).append(t)
        groups = []
        for members in by_key.values():
            for k in range(0, len(members), self.max_batch):
                groups.append(members[k:k + self.max_batch])
        return groups

    def _launch(self, group: list[Ticket]) -> None:
        s = min(range(self.backend.n_streams), key=lambda k: self._stream_load[k])
        now = time.perf_counter()
        for t in group:
            t.t_admit = now
        try:
            if self.batching:
                handle = self.backend.launch_batch([t.request for t in group], s)
            else:
                handle = self.backend.launch(group[0].request, s)
        except BaseException as e:  # noqa: BLE001 - reported on the tickets
            with self._lock:
                self.bytes_in_use -= sum(t.footprint for t in group)
            for t in group:
                t._finish(error=e)
                self.completed.append(t)
            return
        self._stream_load[s] += 1
        self._inflight.append((group, handle, s))

    def step(self) -> int:
        """One scheduling round: retire finished work, admit and launch. Returns the
        number of requests still queued or in flight."""
        still = []
        for group, handle, s in self._inflight:
            try:
                finished = self.backend.done(handle)
            except BaseException as e:  # noqa: BLE001 - reported on the tickets
                self._retire(group, handle, s, error=e)
                continue
            if finished:
                try:
                    if self.batching:
                        values = self.backend.collect_batch([t.request for t in group], handle)
                    else:
                        values = [self.backend.collect(group[0].request, handle)]
                except BaseException as e:  # noqa: BLE001
                    self._retire(group, handle, s, error=e)
                else:
                    self._retire(group, handle, s, values=values)
            else:
                still.append((group, handle, s))
        self._inflight = still
        admitted = self._admit()
        # a launch slot per group: batches of one shape, at most max_launches in flight
        for group in self._groups(admitted):
            self._launch(group)
        with self._lock:
            return len(self._queue) + self._in_flight_requests()

    def _retire(self, group: list[Ticket], handle, s: int, values=None, error=None):
        try:
            self.backend.release(handle)
        finally:
            self._stream_load[s] -= 1
            with self._lock:
                self.bytes_in_use -= sum(t.footprint for t in group)
            now = time.perf_counter()
            for k, t in enumerate(group):
                t._finish(value=None if values is None else values[k], error=error, now=now)
                self.completed.append(t)

    def drain(self, timeout: float = 600.0) -> None:
        """Run the scheduler until nothing is queued or in flight."""
        end = time.perf_counter() + timeout
        while self.step():
            if time.perf_counter() > end:
                raise TimeoutError("server did not drain")
            time.sleep(self.poll_interval)

    def start(self) -> None:
        """Serve from a background thread until :meth:`stop`."""
        if self._thread is not None:
            return
        self._stop.clear()

        def loop():
            while not self._stop.is_set():
                if not self.step():
                    time.sleep(self.poll_interval)

        self._thread = threading.Thread(target=loop, name="ffsim-amd-server", daemon=True)
        self._thread.start()

    def stop(self) -> None:
        if self._thread is None:
            return
        self._stop.set()
        self._thread.join()
        self._thread = None
        self.drain()

    # -- reporting ---------------------------------------------------------
    def stats(self) -> dict:
        ok = [t for t in self.completed if t._error is None]
        lat = np.array([t.latency for t in ok]) if ok else np.zeros(0)
        span = (max(t.t_done for t in ok) - min(t.t_arrival for t in ok)) if ok else 0.0
        return {
            "served": len(ok),
            "failed": len(self.completed) - len(ok),
            "p50_ms": float(np.percentile(lat, 50) * 1e3) if lat.size else None,
            "p99_ms": float(np.percentile(lat, 99) * 1e3) if lat.size else None,
            "states_per_s": (len(ok) / span) if span > 0 else None,
            "peak_hbm_bytes": self.peak_bytes,
            "hbm_budget": self.hbm_budget,
        }
