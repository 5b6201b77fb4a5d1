"""Adjoint differentiation of the spin-balanced UCJ state.

ffsim has no analytic gradient: its optimizers take central finite
differences, two energy evaluations per parameter
(python/ffsim/optimize/_util.py:60-110).  Here the gradient of any real loss
of psi(theta) = UCJ(theta) |psi_0> costs one backward sweep, about twice the
work of one application of the ansatz:

* the state is un-computed gate by gate (every gate is unitary, so no
  intermediate state is stored) while the adjoint state lam = dL/dconj(psi)
  is carried backward through the same gates;
* at each diagonal Coulomb gate the phase backward reduces
  w = Im(conj(lam) psi) to row sums, column sums and the beta-occupation
  weighted rows on the GPU (``ffsim_diag_coulomb_grad_reduce``);
* at each orbital rotation the transition one-body RDM
  rho_pq = <lam| a+_p a_q |psi> (``ffsim_transition_rdm``) is the gradient of
  the rotation in its Lie algebra; the host chains it through the products
  V_k = U_k^dag U_{k-1} and the exponential map U = expm(K(theta))
  (python/ffsim/linalg/util.py:93-124, 151-168) in FP64.

The parameter layout is UCJOpSpinBalanced's
(python/ffsim/variational/ucj_spin_balanced.py:197-293).  The same sweep is
exposed to torch as :class:`UCJStateFunction`, so ``loss.backward()`` works
through :func:`ucj_state`.
"""

from __future__ import annotations

import itertools
import math
from dataclasses import dataclass

import numpy as np
import scipy.linalg
import torch

from ffsim_amd import _cistring, _device, _lib
from ffsim_amd.linalg import antihermitians_from_parameters
from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced


@dataclass(frozen=True)
class UCJLayout:
    """Where each block of the real parameter vector goes."""

    norb: int
    n_reps: int
    pairs_aa: tuple
    pairs_ab: tuple
    with_final_orbital_rotation: bool

    @staticmethod
    def make(norb, n_reps, interaction_pairs=None, with_final_orbital_rotation=False):
        pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
        triu = list(itertools.combinations_with_replacement(range(norb), 2))
        pairs_aa = tuple(triu if pairs_aa is None else pairs_aa)
        pairs_ab = tuple(triu if pairs_ab is None else pairs_ab)
        return UCJLayout(norb, n_reps, pairs_aa, pairs_ab, bool(with_final_orbital_rotation))

    @property
    def n_params(self) -> int:
        per_rep = self.norb**2 + len(self.pairs_aa) + len(self.pairs_ab)
        return self.n_reps * per_rep + self.with_final_orbital_rotation * self.norb**2

    def unpack(self, params: np.ndarray):
        """(K, J, K_final): anti-Hermitian generators, diag-Coulomb matrices."""
        params = np.asarray(params, dtype=float)
        if params.shape != (self.n_params,):
            raise ValueError(
                f"Expected {self.n_params} parameters but got {params.size}."
            )
        n = self.norb
        K = np.zeros((self.n_reps, n, n), dtype=complex)
        J = np.zeros((self.n_reps, 2, n, n))
        index = 0
        for r in range(self.n_reps):
            K[r] = antihermitians_from_parameters(params[index : index + n * n], n, 1)[0]
            index += n * n
            for kind, pairs in enumerate((self.pairs_aa, self.pairs_ab)):
                if pairs:
                    rows, cols = zip(*pairs)
                    vals = params[index : index + len(pairs)]
                    J[r, kind][cols, rows] = vals
                    J[r, kind][rows, cols] = vals
                    index += len(pairs)
        K_final = None
        if self.with_final_orbital_rotation:
            K_final = antihermitians_from_parameters(params[index:], n, 1)[0]
        return K, J, K_final


def _generator_gradient(R: np.ndarray) -> np.ndarray:
    """dL/dparams of K(params) given dL = Re tr(dK R) (layout of linalg/util.py:93-124)."""
    n = R.shape[0]
    n_triu = n * (n - 1) // 2
    out = np.zeros(n * n)
    r, c = np.triu_indices(n, k=1)
    out[:n_triu] = (R[c, r] - R[r, c]).real
    r, c = np.triu_indices(n)
    out[n_triu:] = -(R[c, r] + R[r, c]).imag
    diag = r == c
    out[n_triu:][diag] = -R[r[diag], r[diag]].imag
    return out


def _expm_adjoint(K: np.ndarray, Q: np.ndarray) -> np.ndarray:
    """R with Re tr(dK R) = Re tr((dU U^dag) Q) for U = expm(K), K anti-Hermitian.

    dU U^dag = int_0^1 e^{sK} dK e^{-sK} ds, so R = int_0^1 e^{-sK} Q e^{sK} ds,
    evaluated in the eigenbasis of the Hermitian -iK: entry (a, b) gets the
    factor (e^z - 1)/z with z = i(w_b - w_a).
    """
    w, V = np.linalg.eigh(-1j * K)
    Qt = V.conj().T @ Q @ V
    z = 1j * (w[None, :] - w[:, None])
    small = np.abs(z) < 1e-8
    phi = np.where(small, 1.0 + 0.5 * z, np.expm1(np.where(small, 1.0, z)) / np.where(small, 1.0, z))
    return V @ (Qt * phi) @ V.conj().T


def _pair_gradient(G: np.ndarray, pairs) -> np.ndarray:
    """Gradient of symmetric entries J[p,q] = J[q,p] = theta from dL/dJ (entrywise)."""
    return np.array([G[p, q] + G[q, p] if p != q else G[p, p] for p, q in pairs])


class DeviceAdjointOps:
    """The primitive steps of the sweep on the GPU (states are 1-D device tensors)."""

    def __init__(self, norb: int, nelec):
        self.norb = norb
        self.nelec = tuple(nelec)
        na, nb = self.nelec
        self.dims = (math.comb(norb, na), math.comb(norb, nb))
        occ = []
        for nocc in self.nelec:
            strings = _cistring.make_strings(range(norb), nocc)
            occ.append(((strings[:, None] >> np.arange(norb)) & 1).astype(float))
        self.occ_a, self.occ_b = occ

    def clone(self, t):
        return t.clone()

    def rotate(self, t, U):
        from ffsim_amd.gates.orbital_rotation import apply_orbital_rotation

        return apply_orbital_rotation(t, U, self.norb, self.nelec, copy=False)

    def rotate_both(self, a, b, U):
        """Both states rotated by U in one call (small states share launches)."""
        from ffsim_amd.gates.orbital_rotation import orbital_rotations_in_place

        orbital_rotations_in_place([a, b], [U, U], self.norb, self.nelec)
        return a, b

    def diag(self, t, j_aa, j_ab, time):
        from ffsim_amd.gates.diag_coulomb import apply_diag_coulomb_evolution

        return apply_diag_coulomb_evolution(
            t, (j_aa, j_ab, j_aa), time, self.norb, self.nelec, copy=False
        )

    def diag_both(self, a, b, j_aa, j_ab, time):
        """Both states evolved by the same diagonal gate in one launch."""
        from ffsim_amd.gates.diag_coulomb import diag_coulomb_evolutions_in_place

        diag_coulomb_evolutions_in_place([a, b], (j_aa, j_ab, j_aa), time, self.norb,
                                         self.nelec, False)
        return a, b

    def transition_rdm(self, lam, psi):
        out = torch.empty((2, self.norb, self.norb), dtype=torch.complex128, device=psi.device)
        na, nb = self.nelec
        _lib.call(
            "ffsim_transition_rdm", _device.ptr(lam), _device.ptr(psi), _device.dtype_tag(psi),
            self.norb, na, nb, _device.ptr(out), _device.stream_ptr(),
        )
        return out

    def phase_reductions(self, psi, lam):
        da, db = self.dims
        out = torch.empty(da + db + da * self.norb, dtype=torch.float64, device=psi.device)
        na, nb = self.nelec
        _lib.call(
            "ffsim_diag_coulomb_grad_reduce", _device.ptr(psi), _device.ptr(lam),
            _device.dtype_tag(psi), self.norb, na, nb, _device.ptr(out), _device.stream_ptr(),
        )
        return out

    def host_rdm(self, obj) -> np.ndarray:
        return obj.cpu().numpy()

    def host_phase(self, obj):
        da, db = self.dims
        a = obj.cpu().numpy()
        return a[:da], a[da : da + db], a[da + db :].reshape(da, self.norb)


def ucj_adjoint_sweep(ops, layout: UCJLayout, params: np.ndarray, psi_out, lam):
    """Backward sweep: (dL/dparams, dL/dconj(psi_0)) from psi_out and lam = dL/dconj(psi_out).

    ``ops`` supplies the primitive steps (rotate, diag, transition_rdm,
    phase_reductions and their host reads; optionally rotate_both and
    diag_both, one call that transforms psi and lam alike).  ``psi_out`` and ``lam`` are not
    modified.  The gate list is the merged form of _apply_unitary_
    (ucj_spin_balanced.py:657-696): V_1 = U_1^dag, D_1, V_2 = U_2^dag U_1, ...,
    D_R, V_{R+1} = F U_R.
    """
    norb = layout.norb
    R = layout.n_reps
    K, J, K_final = layout.unpack(params)
    U = np.array([scipy.linalg.expm(k) for k in K]) if R else np.zeros((0, norb, norb))
    F = scipy.linalg.expm(K_final) if K_final is not None else np.eye(norb)
    eye = np.eye(norb, dtype=complex)
    V = []
    prev = eye
    for r in range(R):
        V.append(U[r].conj().T @ prev)
        prev = U[r]
    V.append(F @ prev)

    psi = ops.clone(psi_out)
    lam = ops.clone(lam)
    rdms = [None] * (R + 1)
    phases = [None] * R
    for k in range(R, -1, -1):
        # rotation V_{k+1} (0-based index k), measured after the gate
        rdms[k] = ops.transition_rdm(lam, psi)
        Vh = V[k].conj().T
        if hasattr(ops, "rotate_both"):
            psi, lam = ops.rotate_both(psi, lam, Vh)
        else:
            psi = ops.rotate(psi, Vh)
            lam = ops.rotate(lam, Vh)
        if k > 0:
            j_aa, j_ab = J[k - 1]
            phases[k - 1] = ops.phase_reductions(psi, lam)
            if hasattr(ops, "diag_both"):
                psi, lam = ops.diag_both(psi, lam, j_aa, j_ab, 1.0)
            else:
                psi = ops.diag(psi, j_aa, j_ab, 1.0)
                lam = ops.diag(lam, j_aa, j_ab, 1.0)

    rho = [ops.host_rdm(x).sum(axis=0) for x in rdms]
    grad = np.zeros(layout.n_params)
    index = 0
    n2 = norb * norb
    P = [U[r] @ rho[r].T @ U[r].conj().T for r in range(R)]
    time = -1.0
    for r in range(R):
        if r + 1 < R:
            Q = P[r + 1] - P[r]
        else:
            Q = F.conj().T @ rho[R].T @ F - P[r]
        grad[index : index + n2] = _generator_gradient(_expm_adjoint(K[r], Q))
        index += n2
        rsum, csum, M = ops.host_phase(phases[r])
        occ_a, occ_b = ops.occ_a, ops.occ_b
        g_aa = 0.5 * time * ((occ_a.T * rsum) @ occ_a + (occ_b.T * csum) @ occ_b)
        g_ab = time * (occ_a.T @ M)
        for G, pairs in ((g_aa, layout.pairs_aa), (g_ab, layout.pairs_ab)):
            if pairs:
                grad[index : index + len(pairs)] = _pair_gradient(G, pairs)
                index += len(pairs)
    if K_final is not None:
        grad[index:] = _generator_gradient(_expm_adjoint(K_final, rho[R].T))
    return grad, lam


def _apply_forward(layout: UCJLayout, params: np.ndarray, t: torch.Tensor, nelec) -> torch.Tensor:
    K, J, K_final = layout.unpack(params)
    op = UCJOpSpinBalanced(
        diag_coulomb_mats=J,
        orbital_rotations=np.array([scipy.linalg.expm(k) for k in K]),
        final_orbital_rotation=scipy.linalg.expm(K_final) if K_final is not None else None,
        validate=False,
    )
    return op.apply_in_place(t, layout.norb, nelec)


class UCJStateFunction(torch.autograd.Function):
    """psi = UCJ(params) psi_0 with the adjoint sweep as its backward.

    ``params`` is a real float64 tensor (any device; the chain rule runs on the
    host), ``reference_state`` a complex 1-D tensor on the GPU.
    """

    @staticmethod
    def forward(ctx, params, reference_state, layout, nelec):
        p = params.detach().cpu().numpy().astype(float)
        psi = reference_state.detach().reshape(-1).clone()
        _device.check_size(psi, layout.norb, nelec)
        _apply_forward(layout, p, psi, nelec)
        ctx.layout = layout
        ctx.nelec = tuple(nelec)
        ctx.params_np = p
        ctx.params_meta = (params.device, params.dtype)
        ctx.save_for_backward(psi)
        return psi

    @staticmethod
    def backward(ctx, grad_psi):
        (psi,) = ctx.saved_tensors
        lam = grad_psi.reshape(-1).to(psi.dtype).contiguous()
        ops = DeviceAdjointOps(ctx.layout.norb, ctx.nelec)
        grad, lam0 = ucj_adjoint_sweep(ops, ctx.layout, ctx.params_np, psi, lam)
        dev, dtype = ctx.params_meta
        grad_params = torch.as_tensor(grad, dtype=dtype, device=dev)
        grad_ref = lam0 if ctx.needs_input_grad[1] else None
        return grad_params, grad_ref, None, None


def ucj_state(params, reference_state, *, norb: int, nelec, n_reps: int,
              interaction_pairs=None, with_final_orbital_rotation: bool = False) -> torch.Tensor:
    """Differentiable UCJ state: ``UCJOpSpinBalanced.from_parameters(params, ...)``
    applied to ``reference_state`` (numpy or device tensor), on the GPU."""
    layout = UCJLayout.make(norb, n_reps, interaction_pairs, with_final_orbital_rotation)
    if not isinstance(params, torch.Tensor):
        params = torch.as_tensor(np.asarray(params, dtype=float))
    if not isinstance(reference_state, torch.Tensor) or reference_state.device.type != "cuda":
        reference_state = _device.to_device(np.asarray(reference_state), copy=True).tensor
    return UCJStateFunction.apply(params, reference_state, layout, tuple(nelec))


class HermitianApply(torch.autograd.Function):
    """y = H x for a device linear operator (backward: H^dag g, i.e. H g when Hermitian)."""

    @staticmethod
    def forward(ctx, x, linop):
        ctx.linop = linop
        return linop.matvec_device(x.detach().reshape(-1).clone())

    @staticmethod
    def backward(ctx, g):
        adjoint = getattr(ctx.linop, "rmatvec_device", None) or ctx.linop.matvec_device
        return adjoint(g.reshape(-1).contiguous().clone()), None


def expectation(psi: torch.Tensor, linop) -> torch.Tensor:
    """Re <psi|H|psi> as a differentiable scalar (H Hermitian, e.g. a DF Hamiltonian linop)."""
    return torch.vdot(psi, HermitianApply.apply(psi, linop)).real


def ucj_energy_and_gradient(params, linop, reference_state, *, norb: int, nelec, n_reps: int,
                            interaction_pairs=None, with_final_orbital_rotation: bool = False):
    """Energy <psi(theta)|H|psi(theta)> and its gradient: one forward application,
    one matvec and one adjoint sweep (lam = 2 H psi)."""
    layout = UCJLayout.make(norb, n_reps, interaction_pairs, with_final_orbital_rotation)
    p = np.asarray(params, dtype=float)
    if isinstance(reference_state, torch.Tensor) and reference_state.device.type == "cuda":
        psi = reference_state.reshape(-1).clone()
    else:
        psi = _device.to_device(np.asarray(reference_state), copy=True).tensor
    _apply_forward(layout, p, psi, tuple(nelec))
    h_psi = linop.matvec_device(psi.clone())
    energy = float(torch.vdot(psi, h_psi).real.item())
    ops = DeviceAdjointOps(norb, nelec)
    grad, _ = ucj_adjoint_sweep(ops, layout, p, psi, 2.0 * h_psi)
    return energy, grad
