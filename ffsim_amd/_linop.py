"""scipy LinearOperator whose matvec runs on the GPU.

numpy vectors go through scipy's usual path (copied to HBM and back); torch
tensors on the GPU are handled resident by ``matvec``/``@`` without leaving
the device.  ``matvec == rmatvec`` as in ffsim (the operators are Hermitian).
"""

from __future__ import annotations

import math
from typing import Callable

import numpy as np
import torch
from scipy.sparse.linalg import LinearOperator

from ffsim_amd import _device


def _dim(norb, nelec):
    if isinstance(nelec, int):
        return math.comb(norb, nelec)
    return math.comb(norb, nelec[0]) * math.comb(norb, nelec[1])


class DeviceLinearOperator(LinearOperator):
    """``rmatvec_device`` defaults to ``matvec_device`` (Hermitian operators)."""

    def __init__(self, norb: int, nelec, matvec_device: Callable[[torch.Tensor], torch.Tensor],
                 rmatvec_device: Callable[[torch.Tensor], torch.Tensor] | None = None):
        n = _dim(norb, nelec)
        super().__init__(dtype=np.complex128, shape=(n, n))
        self.norb = norb
        self.nelec = nelec
        self.matvec_device = matvec_device
        self.rmatvec_device = rmatvec_device
        self.hermitian = rmatvec_device is None

    def _apply_host(self, fn, x):
        x = np.asarray(x)
        shape = x.shape
        st = _device.to_device(x.reshape(-1), copy=False)
        out = fn(st.tensor)
        return st.restore(out).reshape(shape)

    def _matvec(self, x):
        return self._apply_host(self.matvec_device, x)

    def _rmatvec(self, x):
        return self._apply_host(self.rmatvec_device or self.matvec_device, x)

    def _adjoint(self):
        if self.hermitian:
            return self
        return DeviceLinearOperator(self.norb, self.nelec, self.rmatvec_device, self.matvec_device)

    def matvec(self, x):
        if isinstance(x, torch.Tensor) and x.device.type == "cuda":
            return self.matvec_device(x.reshape(-1))
        return super().matvec(x)

    def __matmul__(self, x):
        if isinstance(x, torch.Tensor) and x.device.type == "cuda":
            return self.matvec_device(x.reshape(-1))
        return super().__matmul__(x)

    def dot(self, x):
        if isinstance(x, torch.Tensor) and x.device.type == "cuda":
            return self.matvec_device(x.reshape(-1))
        return super().dot(x)
