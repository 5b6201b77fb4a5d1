"""Device-state plumbing: numpy / torch inputs onto HBM and back.

States live in HBM as 1-D torch tensors (complex128 or complex64) viewed by
the kernels as dim_a x dim_b row-major matrices.  numpy inputs are copied to
the device and the result is copied back; torch tensors already on the GPU
stay resident and are returned resident.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import Callable

import numpy as np
import torch

from ffsim_amd import _lib


def require_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError(
            "ffsim_amd gate kernels need a ROCm GPU (torch.cuda.is_available() is False); "
            "there is no CPU fallback"
        )
    return torch.device("cuda", torch.cuda.current_device())


def dtype_tag(t: torch.Tensor) -> int:
    if t.dtype == torch.complex128:
        return _lib.C128
    if t.dtype == torch.complex64:
        return _lib.C64
    raise TypeError(f"state dtype must be complex128 or complex64, got {t.dtype}")


def ptr(t: torch.Tensor | None):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def host_mat(m, shape=None, dtype=np.complex128):
    """C-contiguous host array (or None) for the C ABI."""
    if m is None:
        return None
    a = np.ascontiguousarray(np.asarray(m), dtype=dtype)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def host_ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


# Copies between pageable host memory and the GPU go through the runtime's
# small staging buffers (about 4 GB/s measured for a 188 MB state); through a
# pinned bounce buffer they run at PCIe rate.  States below this size are
# copied directly.  Pinned buffers come from torch's caching host allocator,
# so the pinned footprint stays at the largest state copied.
_STAGE_MIN_BYTES = 1 << 20


# Large host-to-device copies run in chunks through two pinned buffers, so
# that the host copy of one chunk overlaps the transfer of the previous one.
_CHUNK_BYTES = 64 << 20


def host_to_device(arr: np.ndarray, dev: torch.device) -> torch.Tensor:
    """Copy a C-contiguous 1-D host array to a new device tensor (current stream)."""
    src = torch.from_numpy(arr)
    if arr.nbytes < _STAGE_MIN_BYTES:
        return src.to(dev)
    out = torch.empty(src.shape, dtype=src.dtype, device=dev)
    n = src.numel()
    per = max(1, _CHUNK_BYTES // src.element_size())
    bufs = [torch.empty(min(per, n), dtype=src.dtype, pin_memory=True) for _ in range(2)]
    done = [None, None]
    stream = torch.cuda.current_stream(dev)
    for k, i0 in enumerate(range(0, n, per)):
        i1 = min(n, i0 + per)
        slot = k % 2
        if done[slot] is not None:
            done[slot].synchronize()  # the buffer's previous transfer has run
        bufs[slot][: i1 - i0].copy_(src[i0:i1])
        out[i0:i1].copy_(bufs[slot][: i1 - i0], non_blocking=True)
        done[slot] = torch.cuda.Event()
        done[slot].record(stream)
    # the caching host allocator keeps the pinned buffers until their copies have run
    return out


def start_device_to_host(t: torch.Tensor) -> torch.Tensor:
    """Enqueue the copy of a device tensor into pinned host memory on the current
    stream; the result is valid once the stream has reached this point."""
    staged = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    staged.copy_(t, non_blocking=True)
    return staged


def finish_device_to_host(staged: torch.Tensor) -> np.ndarray:
    """Copy a pinned host tensor (its device-to-host copy complete) into a new
    numpy array, so that the pinned buffer returns to the allocator."""
    out = np.empty(tuple(staged.shape), dtype=staged.numpy().dtype)
    torch.from_numpy(out).copy_(staged)
    return out


def device_to_host(t: torch.Tensor) -> np.ndarray:
    """Synchronous copy of a device tensor into a new numpy array.  A large state
    lands in pinned host memory that the returned array keeps (no second host
    copy); the memory goes back to torch's pinned cache when the array is freed."""
    if t.numel() * t.element_size() < _STAGE_MIN_BYTES:
        return t.cpu().numpy()
    staged = start_device_to_host(t)
    torch.cuda.current_stream().synchronize()
    return staged.numpy()


@dataclass
class DeviceState:
    """A working 1-D device tensor plus the conversion back to the caller's type."""

    tensor: torch.Tensor
    restore: Callable[[torch.Tensor], object]


def to_device(vec, copy: bool) -> DeviceState:
    if isinstance(vec, torch.Tensor):
        if vec.device.type != "cuda":
            host = vec.detach().numpy()
            return to_device(host, copy=True)
        t = vec.reshape(-1)
        if t.dtype not in (torch.complex128, torch.complex64):
            t = t.to(torch.complex128)
        elif copy:
            t = t.clone()
        if not t.is_contiguous():
            t = t.contiguous()
        return DeviceState(t, lambda out: out)
    dev = require_device()
    arr = np.asarray(vec)
    if arr.dtype not in (np.complex128, np.complex64):
        arr = arr.astype(np.complex128)
    t = host_to_device(np.ascontiguousarray(arr).reshape(-1), dev)
    return DeviceState(t, device_to_host)


def dims(norb: int, nelec) -> tuple[int, int]:
    if isinstance(nelec, int):
        return math.comb(norb, nelec), 1
    return math.comb(norb, nelec[0]), math.comb(norb, nelec[1])


def check_size(t: torch.Tensor, norb: int, nelec) -> None:
    da, db = dims(norb, nelec)
    if t.numel() != da * db:
        raise ValueError(
            f"state has {t.numel()} entries but norb={norb}, nelec={nelec} needs {da * db}"
        )


def work_elements(t: torch.Tensor, norb: int, na: int, nb: int) -> int:
    """Elements of the transposition buffer an orbital rotation of t needs (0: none)."""
    return _lib.check(
        _lib.lib().ffsim_orbital_rotation_work_elements(dtype_tag(t), norb, na, nb)
    )


def work_buffer(t: torch.Tensor, norb: int, na: int, nb: int) -> torch.Tensor | None:
    n = work_elements(t, norb, na, nb)
    if n == 0:
        return None
    return torch.empty(n, dtype=t.dtype, device=t.device)
