"""Trotter evolution: double-factorized (python/ffsim/trotter/double_factorized.py)
and split-operator for diagonal Coulomb Hamiltonians
(python/ffsim/trotter/diagonal_coulomb_split_op.py).

The host builds the gate sequence of ffsim's schedule (one-body terms merged
into the running basis as ffsim does, trotter/double_factorized.py:89-127;
Suzuki's symmetric recursion for order >= 1, trotter/_util.py:18-55) and the
GPU library applies it as one rotated diag-Coulomb sequence.
"""

from __future__ import annotations

import cmath
from collections.abc import Iterator

import numpy as np
import scipy.linalg

from ffsim_amd import _device, _lib
from ffsim_amd.gates.diag_coulomb import prepared_mats
from ffsim_amd.hamiltonians import DiagonalCoulombHamiltonian, DoubleFactorizedHamiltonian


def simulate_trotter_step_iterator(n_terms: int, time: float, order: int = 0) -> Iterator[tuple[int, float]]:
    """(term index, time) of one Trotter step (order 0: first order;
    order k >= 1: Suzuki's symmetric formula of order 2k)."""
    if order == 0:
        for i in range(n_terms):
            yield i, time
    else:
        yield from _symmetric(n_terms, time, order)


def _symmetric(n_terms: int, time: float, order: int) -> Iterator[tuple[int, float]]:
    # M. Suzuki, J. Math. Phys. 32, 400 (1991)
    if order == 1:
        for i in range(n_terms - 1):
            yield i, time / 2
        yield n_terms - 1, time
        for i in reversed(range(n_terms - 1)):
            yield i, time / 2
    else:
        split = time / (4 - 4 ** (1 / (2 * order - 1)))
        for _ in range(2):
            yield from _symmetric(n_terms, split, order - 1)
        yield from _symmetric(n_terms, time - 4 * split, order - 1)
        for _ in range(2):
            yield from _symmetric(n_terms, split, order - 1)


def trotter_gate_sequence(hamiltonian: DoubleFactorizedHamiltonian, time: float, norb: int,
                          n_steps: int, order: int):
    """(rotations, diag mats, times): rotations[t] precedes diag term t and the
    last rotation is the final basis change."""
    rotations, mats, times = [], [], []
    step_time = time / n_steps
    current = np.eye(norb, dtype=complex)
    n_terms = 1 + len(hamiltonian.diag_coulomb_mats)
    for _ in range(n_steps):
        for term, t in simulate_trotter_step_iterator(n_terms, step_time, order):
            if term == 0:
                current = scipy.linalg.expm(-1j * t * hamiltonian.one_body_tensor) @ current
            else:
                U = np.asarray(hamiltonian.orbital_rotations[term - 1])
                rotations.append(U.T.conj() @ current)
                mats.append(hamiltonian.diag_coulomb_mats[term - 1])
                times.append(t)
                current = U
    rotations.append(current)
    return rotations, mats, times


def simulate_trotter_double_factorized(
    vec,
    hamiltonian: DoubleFactorizedHamiltonian,
    time: float,
    *,
    norb: int,
    nelec: tuple[int, int],
    n_steps: int = 1,
    order: int = 0,
    copy: bool = True,
):
    """Evolve by exp(-i H time) with the Trotter-Suzuki formula
    (python/ffsim/trotter/double_factorized.py:25-86)."""
    if order < 0:
        raise ValueError(f"the Trotter order cannot be negative (got {order})")
    if n_steps < 0:
        raise ValueError(f"the number of Trotter steps cannot be negative (got {n_steps})")
    st = _device.to_device(vec, copy)
    if n_steps == 0:
        return st.restore(st.tensor)
    rotations, mats, times = trotter_gate_sequence(hamiltonian, time, norb, n_steps, order)
    t = _run_sequence(st.tensor, rotations, [np.asarray(m, dtype=float) for m in mats], times,
                      hamiltonian.z_representation, norb, nelec)
    return st.restore(_apply_constant(t, hamiltonian.constant, time))


def _apply_constant(t, constant: float, time: float):
    if constant:
        from ffsim_amd.hamiltonians.double_factorized_hamiltonian import axpby_

        axpby_(t, t, None, cmath.exp(-1j * time * constant))
    return t


def _run_sequence(t, rotations, mats, times, z_representation: bool, norb: int, nelec):
    """One C-ABI call: ``rotations[k]`` then diag-Coulomb term ``k`` at
    ``times[k]`` for every term, then the last rotation."""
    na, nb = nelec
    _device.check_size(t, norb, nelec)
    rot_arr = np.ascontiguousarray(np.stack(rotations), dtype=np.complex128)
    n_terms = len(mats)
    mat_arr = np.zeros((max(n_terms, 1), 3, norb, norb))
    for k, m in enumerate(mats):
        aa, ab, bb = prepared_mats(m, norb)
        mat_arr[k, 0], mat_arr[k, 1], mat_arr[k, 2] = aa, ab, bb
    time_arr = np.ascontiguousarray(np.asarray(times if times else [0.0], dtype=float))
    work = _device.work_buffer(t, norb, na, nb)
    _lib.call(
        "ffsim_apply_diag_coulomb_sequence",
        _device.ptr(t),
        _device.ptr(work),
        _device.dtype_tag(t),
        norb,
        na,
        nb,
        n_terms,
        _device.host_ptr(rot_arr),
        _device.host_ptr(np.ascontiguousarray(mat_arr)),
        _device.host_ptr(time_arr),
        1 if z_representation else 0,
        _device.stream_ptr(),
    )
    return t


def split_op_gate_sequence(hamiltonian, time: float, norb: int, n_steps: int, order: int):
    """(rotations, diag mats, times) of the split-operator schedule
    (trotter/diagonal_coulomb_split_op.py:87-127): the one-body term
    exp(-i t h) is merged into the running basis, which is applied before each
    diagonal Coulomb term (J_aa, J_ab, J_aa) and reset to the identity."""
    rotations, mats, times = [], [], []
    step_time = time / n_steps
    current = np.eye(norb, dtype=complex)
    aa, ab = (np.asarray(m, dtype=float) for m in hamiltonian.diag_coulomb_mats)
    for _ in range(n_steps):
        for term, t in simulate_trotter_step_iterator(2, step_time, order):
            if term == 0:
                current = scipy.linalg.expm(-1j * t * hamiltonian.one_body_tensor) @ current
            else:
                rotations.append(current)
                mats.append((aa, ab, aa))
                times.append(t)
                current = np.eye(norb, dtype=complex)
    rotations.append(current)
    return rotations, mats, times


def simulate_trotter_diag_coulomb_split_op(
    vec,
    hamiltonian: DiagonalCoulombHamiltonian,
    time: float,
    *,
    norb: int,
    nelec: tuple[int, int],
    n_steps: int = 1,
    order: int = 0,
    copy: bool = True,
):
    """Evolve by exp(-i H time) for a diagonal Coulomb Hamiltonian with the
    split-operator Trotter formula
    (python/ffsim/trotter/diagonal_coulomb_split_op.py:25-84), as one gate
    sequence on the GPU."""
    if order < 0:
        raise ValueError(f"the Trotter order cannot be negative (got {order})")
    if n_steps < 0:
        raise ValueError(f"the number of Trotter steps cannot be negative (got {n_steps})")
    st = _device.to_device(vec, copy)
    if n_steps == 0:
        return st.restore(st.tensor)
    rotations, mats, times = split_op_gate_sequence(hamiltonian, time, norb, n_steps, order)
    t = _run_sequence(st.tensor, rotations, mats, times, False, norb, nelec)
    return st.restore(_apply_constant(t, hamiltonian.constant, time))
