"""CPU oracle for the ffsim gate-application hot path.

TEST INFRASTRUCTURE ONLY.  This module is the correctness oracle (and the
host-core baseline) for the ffsim_amd GPU library.  Only the tests, the smoke
check in ``__graft_entry__`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU reference.  The product
never imports it.

Two layers:

1. Restatement of ffsim's own path (qiskit-community/ffsim): the Python
   schedules of ``python/ffsim/gates/*.py``, ``contract/*.py``,
   ``variational/ucj_spin_balanced.py`` and
   ``hamiltonians/double_factorized_hamiltonian.py`` over C restatements of the
   Rust kernels (``ffsim_oracle.c``).  pyscf's integer tables (not installed
   here) are restated from pyscf's published ``cistring`` algorithms.

2. Independent closed forms used to pin layer 1: the compound-matrix (minors)
   action of an orbital rotation, dense one-body operators built from
   creation/annihilation signs, eigenphases computed from occupation numbers,
   and dense Hamiltonians.  These share no code with layer 1.

Every function of layer 1 cites the reference file it follows; the tests in
``tests/test_oracle_*.py`` check each of them against layer 2 and against the
known answers of the reference's own tests.
"""

from __future__ import annotations

import ctypes
import functools
import itertools
import math
import os
import subprocess
from pathlib import Path

import numpy as np
import scipy.linalg

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libffsim_oracle.so"
# Built for the CPU it runs on (-march=native, as the reference's
# .cargo/config.toml:2 builds ffsim); used when FFSIM_ORACLE_NATIVE=1.
_NATIVE_LIB_PATH = _HERE / "libffsim_oracle_native.so"


def _use_native() -> bool:
    return os.environ.get("FFSIM_ORACLE_NATIVE") == "1"


def build(force: bool = False) -> Path:
    """Compile ffsim_oracle.c (recipe: oracle/Makefile): the portable library,
    or with FFSIM_ORACLE_NATIVE=1 the -march=native one for this host."""
    path = _NATIVE_LIB_PATH if _use_native() else _LIB_PATH
    if force or not path.exists():
        target = ["native"] if _use_native() else []
        subprocess.run(["make", "-B", "-C", str(_HERE), *target] if force else
                       ["make", "-C", str(_HERE), *target], check=True, capture_output=True)
    return path


@functools.cache
def _lib() -> ctypes.CDLL:
    path = build()
    lib = ctypes.CDLL(str(path))
    P = ctypes.c_void_p
    L = ctypes.c_long
    D = ctypes.c_double
    I = ctypes.c_int
    lib.oracle_apply_givens_rotation_in_place.argtypes = [P, L, D, D, D, P, P, L]
    lib.oracle_apply_phase_shift_in_place.argtypes = [P, L, D, D, P, L]
    lib.oracle_apply_num_op_sum_evolution_in_place.argtypes = [P, L, L, L, L, P, P, L]
    lib.oracle_apply_diag_coulomb_evolution_num_rep.argtypes = [P, L, L, I, P, P, P, P, L, P, L]
    lib.oracle_apply_diag_coulomb_evolution_z_rep.argtypes = [P, L, L, I, P, P, P, P, P, P, P, P]
    lib.oracle_contract_diag_coulomb_num_rep.argtypes = [P, P, L, L, I, P, P, P, P, L, P, L]
    lib.oracle_contract_diag_coulomb_z_rep.argtypes = [P, P, L, L, I, P, P, P, P, P]
    lib.oracle_contract_num_op_sum_spin_into_buffer.argtypes = [P, P, L, L, L, L, P, P, L]
    lib.oracle_transpose.argtypes = [P, P, L, L]
    lib.oracle_givens_decomposition.argtypes = [P, I, D, P, P, P, P, P]
    lib.oracle_givens_decomposition.restype = I
    lib.oracle_num_threads.restype = I
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def num_threads() -> int:
    return _lib().oracle_num_threads()


# ---------------------------------------------------------------------------
# pyscf cistring tables (restated from pyscf.fci.cistring, pyscf 2.14.0,
# pinned in the reference's uv.lock:2503-2504; called from
# python/ffsim/_cistring.py:21-30).
# ---------------------------------------------------------------------------


def make_strings(orb_list, nelec: int) -> np.ndarray:
    """pyscf cistring.make_strings: recursive enumeration over orb_list.

    Strings are enumerated as: all strings of the first len-1 orbitals, then
    the same with the last orbital added (pyscf's gen_str_iter).
    """
    orb_list = list(orb_list)
    if nelec == 0:
        return np.asarray([0], dtype=np.int64)
    if nelec > len(orb_list):
        return np.asarray([], dtype=np.int64)

    def gen(orbs, k):
        if k == 1:
            return [1 << o for o in orbs]
        if k >= len(orbs):
            n = 0
            for o in orbs:
                n |= 1 << o
            return [n]
        rest = orbs[:-1]
        this = 1 << orbs[-1]
        res = gen(rest, k)
        for n in gen(rest, k - 1):
            res.append(n | this)
        return res

    return np.asarray(gen(orb_list, nelec), dtype=np.int64)


def gen_occslst(orb_list, nelec: int) -> np.ndarray:
    """pyscf cistring.gen_occslst cast to np.uint (python/ffsim/_cistring.py:27-30)."""
    orb_list = list(orb_list)
    if nelec == 0:
        return np.zeros((1, 0), dtype=np.uint)
    if nelec > len(orb_list):
        return np.zeros((0, nelec), dtype=np.uint)
    strings = make_strings(orb_list, nelec)
    out = np.zeros((len(strings), nelec), dtype=np.uint)
    for k, s in enumerate(strings):
        out[k] = [o for o in range(max(orb_list) + 1) if (int(s) >> o) & 1]
    return out


@functools.cache
def _shifted_orbitals(norb: int, target_orbs: tuple[int, ...]) -> np.ndarray:
    """python/ffsim/gates/orbital_rotation.py:229-236."""
    orbitals = np.arange(norb - len(target_orbs))
    values = sorted(zip(target_orbs, range(norb - len(target_orbs), norb)))
    for index, val in values:
        orbitals = np.insert(orbitals, index, val)
    return orbitals


@functools.cache
def zero_one_subspace_indices(norb: int, nocc: int, target_orbs: tuple[int, int]) -> np.ndarray:
    """python/ffsim/gates/orbital_rotation.py:202-212."""
    orbitals = _shifted_orbitals(norb, target_orbs)
    strings = make_strings(orbitals, nocc)
    indices = np.argsort(strings)
    n00 = math.comb(norb - 2, nocc)
    n11 = math.comb(norb - 2, nocc - 2) if nocc >= 2 else 0
    return indices[n00 : len(indices) - n11].astype(np.uint, copy=False)


@functools.cache
def one_subspace_indices(norb: int, nocc: int, target_orbs: tuple[int, ...]) -> np.ndarray:
    """python/ffsim/gates/orbital_rotation.py:215-226."""
    orbitals = _shifted_orbitals(norb, target_orbs)
    strings = make_strings(orbitals, nocc)
    indices = np.argsort(strings)
    n0 = math.comb(norb, nocc)
    if nocc >= len(target_orbs):
        n0 -= math.comb(norb - len(target_orbs), nocc - len(target_orbs))
    return indices[n0:].astype(np.uint, copy=False)


@functools.cache
def cached_strings(norb: int, nocc: int) -> np.ndarray:
    return make_strings(range(norb), nocc)


@functools.cache
def cached_occslst(norb: int, nocc: int) -> np.ndarray:
    return np.ascontiguousarray(gen_occslst(range(norb), nocc)) if norb > 0 else np.zeros(
        (1 if nocc == 0 else 0, nocc), dtype=np.uint
    )


def init_cache(norb: int, nelec: tuple[int, int]) -> None:
    """Pre-build the tables a gate needs (python/ffsim/init_cache.py:25-41)."""
    for nocc in set(nelec):
        cached_strings(norb, nocc)
        cached_occslst(norb, nocc)
        for i in range(norb - 1):
            zero_one_subspace_indices(norb, nocc, (i, i + 1))
            zero_one_subspace_indices(norb, nocc, (i + 1, i))
        for i in range(norb):
            one_subspace_indices(norb, nocc, (i,))


# ---------------------------------------------------------------------------
# Givens decomposition (src/linalg/givens.rs:69-149 restated in C)
# ---------------------------------------------------------------------------


def givens_decomposition(mat: np.ndarray, tol: float = 1e-12):
    mat = np.ascontiguousarray(mat, dtype=np.complex128)
    n = mat.shape[0]
    if mat.ndim != 2 or mat.shape[0] != mat.shape[1]:
        raise ValueError("mat must be a square matrix")
    cap = max(n * (n - 1) // 2, 1)
    c = np.zeros(cap)
    s = np.zeros(2 * cap)
    i = np.zeros(cap, dtype=np.int32)
    j = np.zeros(cap, dtype=np.int32)
    d = np.zeros(2 * max(n, 1))
    nr = _lib().oracle_givens_decomposition(_p(mat), n, tol, _p(c), _p(s), _p(i), _p(j), _p(d))
    rots = [(float(c[k]), complex(s[2 * k], s[2 * k + 1]), int(i[k]), int(j[k])) for k in range(nr)]
    diag = d[: 2 * n].view(np.complex128).copy()
    return rots, diag


# ---------------------------------------------------------------------------
# Gate schedules (python/ffsim/gates/*.py) over the C kernels
# ---------------------------------------------------------------------------


def _givens_in_place(vec2d: np.ndarray, c: float, s: complex, slice1, slice2) -> None:
    """_lib.apply_givens_rotation_in_place (src/gates/orbital_rotation.rs:19-104)."""
    assert vec2d.flags.c_contiguous and vec2d.dtype == np.complex128
    slice1 = np.ascontiguousarray(slice1, dtype=np.uint64)
    slice2 = np.ascontiguousarray(slice2, dtype=np.uint64)
    _lib().oracle_apply_givens_rotation_in_place(
        _p(vec2d), vec2d.shape[1], c, s.real, s.imag, _p(slice1), _p(slice2), len(slice1)
    )


def _phase_shift_in_place(vec2d: np.ndarray, phase: complex, indices) -> None:
    """_lib.apply_phase_shift_in_place (src/gates/phase_shift.rs:16-30)."""
    indices = np.ascontiguousarray(indices, dtype=np.uint64)
    _lib().oracle_apply_phase_shift_in_place(
        _p(vec2d), vec2d.shape[1], phase.real, phase.imag, _p(indices), len(indices)
    )


def _transpose_copy(vec2d: np.ndarray) -> np.ndarray:
    """np.ascontiguousarray(vec.T) (python/ffsim/gates/orbital_rotation.py:143)."""
    out = np.empty((vec2d.shape[1], vec2d.shape[0]), dtype=np.complex128)
    _lib().oracle_transpose(_p(vec2d), _p(out), vec2d.shape[0], vec2d.shape[1])
    return out


def _get_givens_decomposition(mat):
    """python/ffsim/gates/orbital_rotation.py:157-174."""
    if isinstance(mat, np.ndarray) and mat.ndim == 2:
        decomp = givens_decomposition(mat)
        return decomp, decomp
    mat_a, mat_b = mat
    return (
        givens_decomposition(mat_a) if mat_a is not None else None,
        givens_decomposition(mat_b) if mat_b is not None else None,
    )


def _apply_adjacent(vec2d, c, s, target_orbs, norb, nocc) -> None:
    """python/ffsim/gates/orbital_rotation.py:177-199."""
    i, j = target_orbs
    assert i == j + 1 or i == j - 1, "Target orbitals must be adjacent."
    indices = zero_one_subspace_indices(norb, nocc, target_orbs)
    half = len(indices) // 2
    _givens_in_place(vec2d, c, s, indices[:half], indices[half:])


def apply_orbital_rotation(vec, mat, norb, nelec, *, copy=True):
    """ffsim.apply_orbital_rotation (python/ffsim/gates/orbital_rotation.py:44-154)."""
    vec = np.asarray(vec, dtype=np.complex128)
    if copy:
        vec = vec.copy()
    if isinstance(nelec, int):
        rots, phases = givens_decomposition(mat)
        v = np.ascontiguousarray(vec.reshape((-1, 1)))
        for c, s, i, j in rots:
            _apply_adjacent(v, c, s.conjugate(), (i, j), norb, nelec)
        for i, ph in enumerate(phases):
            _phase_shift_in_place(v, complex(ph), one_subspace_indices(norb, nelec, (i,)))
        return v.reshape(-1)
    decomp_a, decomp_b = _get_givens_decomposition(mat)
    n_alpha, n_beta = nelec
    dim_a, dim_b = math.comb(norb, n_alpha), math.comb(norb, n_beta)
    v = np.ascontiguousarray(vec.reshape((dim_a, dim_b)))
    if decomp_a is not None:
        rots, phases = decomp_a
        for c, s, i, j in rots:
            _apply_adjacent(v, c, s.conjugate(), (i, j), norb, n_alpha)
        for i, ph in enumerate(phases):
            _phase_shift_in_place(v, complex(ph), one_subspace_indices(norb, n_alpha, (i,)))
    if decomp_b is not None:
        v = _transpose_copy(v)
        rots, phases = decomp_b
        for c, s, i, j in rots:
            _apply_adjacent(v, c, s.conjugate(), (i, j), norb, n_beta)
        for i, ph in enumerate(phases):
            _phase_shift_in_place(v, complex(ph), one_subspace_indices(norb, n_beta, (i,)))
        v = _transpose_copy(v)
    return v.reshape(-1)


def conjugate_orbital_rotation(orbital_rotation):
    """python/ffsim/contract/diag_coulomb.py:28-38."""
    if isinstance(orbital_rotation, np.ndarray) and orbital_rotation.ndim == 2:
        return orbital_rotation.T.conj()
    a, b = orbital_rotation
    return (a.T.conj() if a is not None else None, b.T.conj() if b is not None else None)


def get_mat_exp(mat, time, norb, z_representation):
    """python/ffsim/gates/diag_coulomb.py:223-275."""
    if isinstance(mat, np.ndarray) and mat.ndim == 2:
        mat_aa, mat_ab = mat.copy(), mat.copy()
        mat_aa[np.diag_indices(norb)] *= 0.5
        if z_representation:
            mat_aa *= 0.25
            mat_ab *= 0.25
        e_aa = np.exp(-1j * time * mat_aa)
        e_ab = np.exp(-1j * time * mat_ab)
        return e_aa, e_ab, e_aa
    mat_aa, mat_ab, mat_bb = mat
    out = []
    for m, halve in ((mat_aa, True), (mat_ab, False), (mat_bb, True)):
        if m is None:
            out.append(np.ones((norb, norb), dtype=complex))
            continue
        m = np.array(m, dtype=float, copy=True)
        if halve:
            m[np.diag_indices(norb)] *= 0.5
        if z_representation:
            m *= 0.25
        out.append(np.exp(-1j * time * m))
    return out[0], out[1], out[2]


def apply_diag_coulomb_evolution(
    vec, mat, time, norb, nelec, *, orbital_rotation=None, z_representation=False, copy=True
):
    """ffsim.apply_diag_coulomb_evolution (python/ffsim/gates/diag_coulomb.py:68-220)."""
    vec = np.asarray(vec, dtype=np.complex128)
    if copy:
        vec = vec.copy()
    if isinstance(nelec, int):
        if z_representation:
            raise NotImplementedError
        nelec = (nelec, 0)
    e_aa, e_ab, e_bb = (np.ascontiguousarray(x) for x in get_mat_exp(mat, time, norb, z_representation))
    n_alpha, n_beta = nelec
    dim_a, dim_b = math.comb(norb, n_alpha), math.comb(norb, n_beta)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, conjugate_orbital_rotation(orbital_rotation), norb, nelec, copy=False)
    v = np.ascontiguousarray(vec.reshape((dim_a, dim_b)))
    if z_representation:
        sa = np.ascontiguousarray(cached_strings(norb, n_alpha))
        sb = np.ascontiguousarray(cached_strings(norb, n_beta))
        conj = [np.ascontiguousarray(x.conj()) for x in (e_aa, e_ab, e_bb)]
        _lib().oracle_apply_diag_coulomb_evolution_z_rep(
            _p(v), dim_a, dim_b, norb, _p(e_aa), _p(e_ab), _p(e_bb),
            _p(conj[0]), _p(conj[1]), _p(conj[2]), _p(sa), _p(sb),
        )
    else:
        oa = np.ascontiguousarray(cached_occslst(norb, n_alpha), dtype=np.uint64)
        ob = np.ascontiguousarray(cached_occslst(norb, n_beta), dtype=np.uint64)
        _lib().oracle_apply_diag_coulomb_evolution_num_rep(
            _p(v), dim_a, dim_b, norb, _p(e_aa), _p(e_ab), _p(e_bb), _p(oa), n_alpha, _p(ob), n_beta
        )
    vec = v.reshape(-1)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, orbital_rotation, norb, nelec, copy=False)
    return vec


def _num_op_in_place(v2d_view, phases, occ) -> None:
    """_lib.apply_num_op_sum_evolution_in_place on a (possibly transposed) view."""
    phases = np.ascontiguousarray(phases, dtype=np.complex128)
    occ = np.ascontiguousarray(occ, dtype=np.uint64)
    rs = v2d_view.strides[0] // 16
    cs = v2d_view.strides[1] // 16
    base = v2d_view if v2d_view.flags.c_contiguous else v2d_view.T
    _lib().oracle_apply_num_op_sum_evolution_in_place(
        _p(base), v2d_view.shape[0], v2d_view.shape[1], rs, cs, _p(phases), _p(occ), occ.shape[1]
    )


def apply_num_op_sum_evolution(vec, coeffs, time, norb, nelec, *, orbital_rotation=None, copy=True):
    """ffsim.apply_num_op_sum_evolution (python/ffsim/gates/num_op_sum.py:62-236)."""
    vec = np.asarray(vec, dtype=np.complex128)
    if copy:
        vec = vec.copy()
    if isinstance(nelec, int):
        phases = np.exp(-1j * time * coeffs)
        occ = cached_occslst(norb, nelec)
        if orbital_rotation is not None:
            vec = apply_orbital_rotation(vec, orbital_rotation.T.conj(), norb, nelec, copy=False)
        v = np.ascontiguousarray(vec.reshape((-1, 1)))
        _num_op_in_place(v, phases, occ)
        vec = v.reshape(-1)
        if orbital_rotation is not None:
            vec = apply_orbital_rotation(vec, orbital_rotation, norb, nelec, copy=False)
        return vec
    if isinstance(coeffs, np.ndarray):
        pa = pb = np.exp(-1j * time * coeffs)
    else:
        ca, cb = coeffs
        pa = np.exp(-1j * time * ca) if ca is not None else None
        pb = np.exp(-1j * time * cb) if cb is not None else None
    n_alpha, n_beta = nelec
    dim_a, dim_b = math.comb(norb, n_alpha), math.comb(norb, n_beta)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, conjugate_orbital_rotation(orbital_rotation), norb, nelec, copy=False)
    v = np.ascontiguousarray(vec.reshape((dim_a, dim_b)))
    if pa is not None:
        _num_op_in_place(v, pa, cached_occslst(norb, n_alpha))
    if pb is not None:
        _num_op_in_place(v.T, pb, cached_occslst(norb, n_beta))
    vec = v.reshape(-1)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, orbital_rotation, norb, nelec, copy=False)
    return vec


# ---------------------------------------------------------------------------
# Contractions and linear operators (python/ffsim/contract/*.py)
# ---------------------------------------------------------------------------


def get_mats(mat, norb, z_representation):
    """python/ffsim/contract/diag_coulomb.py:102-130."""
    if isinstance(mat, np.ndarray):
        mat_aa, mat_ab = mat, mat
        if not z_representation:
            mat_aa = mat_aa.copy()
            mat_aa[np.diag_indices(norb)] *= 0.5
        return mat_aa, mat_ab, mat_aa
    mat_aa, mat_ab, mat_bb = mat
    if mat_aa is None:
        mat_aa = np.zeros((norb, norb))
    elif not z_representation:
        mat_aa = mat_aa.copy()
        mat_aa[np.diag_indices(norb)] *= 0.5
    if mat_bb is None:
        mat_bb = np.zeros((norb, norb))
    elif not z_representation:
        mat_bb = mat_bb.copy()
        mat_bb[np.diag_indices(norb)] *= 0.5
    if mat_ab is None:
        mat_ab = np.zeros((norb, norb))
    return mat_aa, mat_ab, mat_bb


def contract_diag_coulomb(vec, mat, norb, nelec, *, z_representation=False):
    """ffsim.contract.contract_diag_coulomb (python/ffsim/contract/diag_coulomb.py:41-192)."""
    aa, ab, bb = (np.ascontiguousarray(x, dtype=float) for x in get_mats(mat, norb, z_representation))
    vec = np.asarray(vec).astype(complex, copy=False)
    n_alpha, n_beta = nelec
    dim_a, dim_b = math.comb(norb, n_alpha), math.comb(norb, n_beta)
    v = np.ascontiguousarray(vec.reshape((dim_a, dim_b)))
    out = np.zeros_like(v)
    if z_representation:
        sa = np.ascontiguousarray(cached_strings(norb, n_alpha))
        sb = np.ascontiguousarray(cached_strings(norb, n_beta))
        _lib().oracle_contract_diag_coulomb_z_rep(
            _p(v), _p(out), dim_a, dim_b, norb, _p(aa), _p(ab), _p(bb), _p(sa), _p(sb)
        )
    else:
        oa = np.ascontiguousarray(cached_occslst(norb, n_alpha), dtype=np.uint64)
        ob = np.ascontiguousarray(cached_occslst(norb, n_beta), dtype=np.uint64)
        _lib().oracle_contract_diag_coulomb_num_rep(
            _p(v), _p(out), dim_a, dim_b, norb, _p(aa), _p(ab), _p(bb), _p(oa), n_alpha, _p(ob), n_beta
        )
    return out.reshape(-1)


def diag_coulomb_matvec(vec, mat, norb, nelec, *, orbital_rotation=None, z_representation=False):
    """matvec of ffsim.contract.diag_coulomb_linop (python/ffsim/contract/diag_coulomb.py:245-268)."""
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, conjugate_orbital_rotation(orbital_rotation), norb, nelec)
    vec = contract_diag_coulomb(vec, mat, norb, nelec, z_representation=z_representation)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, orbital_rotation, norb, nelec, copy=False)
    return vec


def contract_num_op_sum(vec, coeffs, norb, nelec):
    """ffsim.contract.contract_num_op_sum (python/ffsim/contract/num_op_sum.py:27-72)."""
    vec = np.asarray(vec).astype(complex, copy=False)
    coeffs = np.ascontiguousarray(coeffs, dtype=float)
    n_alpha, n_beta = nelec
    dim_a, dim_b = math.comb(norb, n_alpha), math.comb(norb, n_beta)
    v = np.ascontiguousarray(vec.reshape((dim_a, dim_b)))
    out = np.zeros_like(v)
    oa = np.ascontiguousarray(cached_occslst(norb, n_alpha), dtype=np.uint64)
    ob = np.ascontiguousarray(cached_occslst(norb, n_beta), dtype=np.uint64)
    lib = _lib()
    lib.oracle_contract_num_op_sum_spin_into_buffer(_p(v), _p(out), dim_a, dim_b, dim_b, 1, _p(coeffs), _p(oa), n_alpha)
    lib.oracle_contract_num_op_sum_spin_into_buffer(_p(v), _p(out), dim_b, dim_a, 1, dim_b, _p(coeffs), _p(ob), n_beta)
    return out.reshape(-1)


def num_op_sum_matvec(vec, coeffs, norb, nelec, *, orbital_rotation=None):
    """matvec of ffsim.contract.num_op_sum_linop (python/ffsim/contract/num_op_sum.py:108-126)."""
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, orbital_rotation.T.conj(), norb, nelec)
    vec = contract_num_op_sum(vec, coeffs, norb, nelec)
    if orbital_rotation is not None:
        vec = apply_orbital_rotation(vec, orbital_rotation, norb, nelec, copy=False)
    return vec


def df_hamiltonian_matvec(vec, one_body_tensor, diag_coulomb_mats, orbital_rotations, constant,
                          norb, nelec, z_representation=False):
    """DoubleFactorizedHamiltonian._linear_operator_ matvec
    (python/ffsim/hamiltonians/double_factorized_hamiltonian.py:244-275)."""
    eigs, vecs = scipy.linalg.eigh(one_body_tensor)
    vec = np.asarray(vec).astype(complex, copy=False)
    result = constant * vec
    result = result + num_op_sum_matvec(vec, eigs, norb, nelec, orbital_rotation=vecs)
    for mat, rot in zip(diag_coulomb_mats, orbital_rotations):
        result = result + diag_coulomb_matvec(vec, mat, norb, nelec, orbital_rotation=rot,
                                              z_representation=z_representation)
    return result


def diagonal_coulomb_hamiltonian_matvec(vec, one_body_tensor, diag_coulomb_mats, constant, norb, nelec):
    """DiagonalCoulombHamiltonian._linear_operator_ matvec
    (python/ffsim/hamiltonians/diagonal_coulomb_hamiltonian.py:68-95)."""
    eigs, vecs = scipy.linalg.eigh(one_body_tensor)
    vec = np.asarray(vec).astype(complex, copy=False)
    result = constant * vec
    result = result + num_op_sum_matvec(vec, eigs, norb, nelec, orbital_rotation=vecs)
    mats = (diag_coulomb_mats[0], diag_coulomb_mats[1], diag_coulomb_mats[0])
    result = result + diag_coulomb_matvec(vec, mats, norb, nelec)
    return result


# ---------------------------------------------------------------------------
# UCJ (python/ffsim/variational/ucj_spin_balanced.py, python/ffsim/linalg/util.py)
# ---------------------------------------------------------------------------


def antihermitian_from_parameters(params: np.ndarray, dim: int) -> np.ndarray:
    """python/ffsim/linalg/util.py:93-124 (one complex matrix)."""
    n_triu = dim * (dim - 1) // 2
    mat = np.zeros((dim, dim), dtype=complex)
    rows, cols = np.triu_indices(dim)
    vals = 1j * params[n_triu:]
    mat[rows, cols] = vals
    mat[cols, rows] = vals
    rows, cols = np.triu_indices(dim, k=1)
    mat[rows, cols] += params[:n_triu]
    mat[cols, rows] -= params[:n_triu]
    return mat


def unitary_from_parameters(params: np.ndarray, dim: int) -> np.ndarray:
    """python/ffsim/linalg/util.py:151-168, 212-232."""
    return scipy.linalg.expm(antihermitian_from_parameters(params, dim))


def ucj_n_params(norb, n_reps, interaction_pairs=None, with_final_orbital_rotation=False) -> int:
    """UCJOpSpinBalanced.n_params (python/ffsim/variational/ucj_spin_balanced.py:143-195)."""
    pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
    n_triu = norb * (norb + 1) // 2
    n_aa = n_triu if pairs_aa is None else len(pairs_aa)
    n_ab = n_triu if pairs_ab is None else len(pairs_ab)
    return n_reps * (n_aa + n_ab + norb**2) + with_final_orbital_rotation * norb**2


def ucj_from_parameters(params, norb, n_reps, interaction_pairs=None, with_final_orbital_rotation=False):
    """UCJOpSpinBalanced.from_parameters (ucj_spin_balanced.py:197-293) -> (J, U, final)."""
    pairs_aa, pairs_ab = interaction_pairs if interaction_pairs is not None else (None, None)
    triu = list(itertools.combinations_with_replacement(range(norb), 2))
    pairs_aa = triu if pairs_aa is None else pairs_aa
    pairs_ab = triu if pairs_ab is None else pairs_ab
    J = np.zeros((n_reps, 2, norb, norb))
    U = np.zeros((n_reps, norb, norb), dtype=complex)
    index = 0
    for r in range(n_reps):
        U[r] = unitary_from_parameters(params[index : index + norb**2], norb)
        index += norb**2
        for t, pairs in enumerate((pairs_aa, pairs_ab)):
            if pairs:
                rows, cols = zip(*pairs)
                vals = params[index : index + len(pairs)]
                J[r, t][cols, rows] = vals
                J[r, t][rows, cols] = vals
                index += len(pairs)
    final = unitary_from_parameters(params[index:], norb) if with_final_orbital_rotation else None
    return J, U, final


def apply_ucj_spin_balanced(vec, diag_coulomb_mats, orbital_rotations, final_orbital_rotation,
                            norb, nelec, copy=True):
    """UCJOpSpinBalanced._apply_unitary_ (python/ffsim/variational/ucj_spin_balanced.py:657-696)."""
    vec = np.asarray(vec, dtype=np.complex128)
    if copy:
        vec = vec.copy()
    current = np.eye(norb)
    for (j_aa, j_ab), U in zip(diag_coulomb_mats, orbital_rotations):
        vec = apply_orbital_rotation(vec, U.T.conj() @ current, norb, nelec, copy=False)
        vec = apply_diag_coulomb_evolution(vec, (j_aa, j_ab, j_aa), -1.0, norb, nelec, copy=False)
        current = U
    last = current if final_orbital_rotation is None else final_orbital_rotation @ current
    return apply_orbital_rotation(vec, last, norb, nelec, copy=False)


def apply_ucj_spin_unbalanced(vec, diag_coulomb_mats, orbital_rotations, final_orbital_rotation,
                              norb, nelec):
    """UCJOpSpinUnbalanced._apply_unitary_ (python/ffsim/variational/ucj_spin_unbalanced.py:705-745)."""
    vec = np.array(vec, dtype=np.complex128)
    eye = np.eye(norb)
    current = np.stack([eye, eye]).astype(complex)
    for mats, rots in zip(diag_coulomb_mats, orbital_rotations):
        step = rots.transpose(0, 2, 1).conj() @ current
        vec = apply_orbital_rotation(vec, (step[0], step[1]), norb, nelec, copy=False)
        vec = apply_diag_coulomb_evolution(vec, (mats[0], mats[1], mats[2]), -1.0, norb, nelec,
                                           copy=False)
        current = rots
    last = current if final_orbital_rotation is None else final_orbital_rotation @ current
    return apply_orbital_rotation(vec, (last[0], last[1]), norb, nelec, copy=False)


def apply_ucj_spinless(vec, diag_coulomb_mats, orbital_rotations, final_orbital_rotation, norb,
                       nelec):
    """UCJOpSpinless._apply_unitary_ (python/ffsim/variational/ucj_spinless.py:456-515)."""
    vec = np.array(vec, dtype=np.complex128)
    current = np.eye(norb, dtype=complex)
    zero = np.zeros((norb, norb))
    for mat, rot in zip(diag_coulomb_mats, orbital_rotations):
        vec = apply_orbital_rotation(vec, rot.T.conj() @ current, norb, nelec, copy=False)
        dc = mat if isinstance(nelec, int) else (mat, zero, mat)
        vec = apply_diag_coulomb_evolution(vec, dc, -1.0, norb, nelec, copy=False)
        current = rot
    last = current if final_orbital_rotation is None else final_orbital_rotation @ current
    return apply_orbital_rotation(vec, last, norb, nelec, copy=False)


def ucj_energy(params, hamiltonian_matvec, reference_state, norb, nelec, n_reps,
               interaction_pairs=None, with_final_orbital_rotation=False) -> float:
    """<psi(theta)|H|psi(theta)> for the spin-balanced UCJ ansatz on the CPU path."""
    J, U, final = ucj_from_parameters(params, norb, n_reps, interaction_pairs,
                                      with_final_orbital_rotation)
    psi = apply_ucj_spin_balanced(reference_state, J, U, final, norb, nelec)
    return float(np.vdot(psi, hamiltonian_matvec(psi)).real)


def gradient_finite_diff(func, theta: np.ndarray, index: int, epsilon: float):
    """Central difference of one parameter (python/ffsim/optimize/_util.py:60-83)."""
    unit = np.zeros(len(theta))
    unit[index] = 1.0
    return (func(theta + epsilon * unit) - func(theta - epsilon * unit)) / (2 * epsilon)


def fd_gradient(func, theta: np.ndarray, epsilon: float = 1e-6, indices=None) -> np.ndarray:
    """Central finite-difference gradient over ``indices`` (all by default), the
    reference's only gradient mechanism (python/ffsim/optimize/_util.py:60-110)."""
    theta = np.asarray(theta, dtype=float)
    indices = range(len(theta)) if indices is None else indices
    return np.array([gradient_finite_diff(func, theta, i, epsilon) for i in indices])


def transition_rdm(lam, psi, norb, nelec) -> np.ndarray:
    """rdm[sigma, p, q] = <lam| a+_{p sigma} a_{q sigma} |psi> by string loops.

    Signs follow the Jordan-Wigner ordering of ascending orbitals within a spin
    (the convention of the orbital-rotation closed form).  Pinned in the tests
    against the dense one-body matrices and against the derivative of the
    minors closed form.
    """
    na, nb = nelec
    sa, sb = colex_strings(norb, na), colex_strings(norb, nb)
    lam = np.asarray(lam).reshape(len(sa), len(sb))
    psi = np.asarray(psi).reshape(len(sa), len(sb))
    out = np.zeros((2, norb, norb), dtype=complex)
    for sigma, strings in enumerate((sa, sb)):
        index = {s: k for k, s in enumerate(strings)}
        for k, s in enumerate(strings):
            for q in range(norb):
                if not (s >> q) & 1:
                    continue
                t0 = s & ~(1 << q)
                sq = bin(s & ((1 << q) - 1)).count("1")
                for p in range(norb):
                    if p != q and (s >> p) & 1:
                        continue
                    sign = (-1) ** (sq + bin(t0 & ((1 << p) - 1)).count("1"))
                    t = index[t0 | (1 << p)]
                    if sigma == 0:
                        val = np.vdot(lam[t], psi[k])
                    else:
                        val = np.vdot(lam[:, t], psi[:, k])
                    out[sigma, p, q] += sign * val
    return out


def cre_des_sign(p: int, q: int, string: int) -> int:
    """Sign of a+_p a_q on a string: parity of the occupied orbitals strictly
    between p and q (pyscf.fci.cistring.cre_des_sign, pyscf 2.14.0)."""
    if p == q:
        return 1
    lo, hi = min(p, q), max(p, q)
    mask = ((1 << hi) - 1) ^ ((1 << (lo + 1)) - 1)
    return -1 if bin(string & mask).count("1") % 2 else 1


def gen_linkstr_index(norb: int, nocc: int) -> np.ndarray:
    """pyscf.fci.cistring.gen_linkstr_index (pyscf 2.14.0, cached by
    python/ffsim/_cistring.py:33-36): per string, rows (a, i, address, sign),
    first the occupied (i, i) entries, then every occupied i to virtual a."""
    strings = cached_strings(norb, nocc)
    index = {int(s): k for k, s in enumerate(strings)}
    n_vir = norb - nocc
    out = np.zeros((len(strings), nocc + nocc * n_vir, 4), dtype=np.int64)
    for k, s in enumerate(strings):
        s = int(s)
        occ = [o for o in range(norb) if (s >> o) & 1]
        vir = [o for o in range(norb) if not (s >> o) & 1]
        row = 0
        for i in occ:
            out[k, row] = (i, i, k, 1)
            row += 1
        for i in occ:
            for a in vir:
                t = s ^ (1 << i) ^ (1 << a)
                out[k, row] = (a, i, index[t], cre_des_sign(a, i, s))
                row += 1
    return out


def contract_one_body(vec, mat, norb, nelec):
    """python/ffsim/contract/one_body.py:24-56: sum_pq M_pq a+_p a_q on both spins,
    as pyscf's contract_1e does it: for every source string and link (a, i, t, sign),
    out[t] += sign * M[a, i] * vec[source]."""
    na, nb = nelec
    da, db = math.comb(norb, na), math.comb(norb, nb)
    v = np.asarray(vec, dtype=np.complex128).reshape(da, db)
    mat = np.asarray(mat, dtype=np.complex128)
    out = np.zeros_like(v)
    for src, links in enumerate(gen_linkstr_index(norb, na)):
        for a, i, t, sign in links:
            out[t] += sign * mat[a, i] * v[src]
    for src, links in enumerate(gen_linkstr_index(norb, nb)):
        for a, i, t, sign in links:
            out[:, t] += sign * mat[a, i] * v[:, src]
    return out.reshape(-1)


def diag_coulomb_phase_reductions(psi, lam, norb, nelec):
    """Row sums, column sums and occupation-weighted row sums of
    w = Im(conj(lam) psi) over the dim_a x dim_b grid."""
    na, nb = nelec
    da, db = math.comb(norb, na), math.comb(norb, nb)
    w = (np.conj(np.asarray(lam)) * np.asarray(psi)).imag.reshape(da, db)
    occ_b = occupations(colex_strings(norb, nb), norb)
    return w.sum(axis=1), w.sum(axis=0), w @ occ_b


def trotter_step_iterator(n_terms, time, order):
    """python/ffsim/trotter/_util.py:18-55."""
    if order == 0:
        for i in range(n_terms):
            yield i, time
        return
    if order == 1:
        for i in range(n_terms - 1):
            yield i, time / 2
        yield n_terms - 1, time
        for i in reversed(range(n_terms - 1)):
            yield i, time / 2
        return
    split = time / (4 - 4 ** (1 / (2 * order - 1)))
    for _ in range(2):
        yield from trotter_step_iterator(n_terms, split, order - 1)
    yield from trotter_step_iterator(n_terms, time - 4 * split, order - 1)
    for _ in range(2):
        yield from trotter_step_iterator(n_terms, split, order - 1)


def simulate_trotter_double_factorized(vec, one_body_tensor, diag_coulomb_mats, orbital_rotations,
                                       constant, time, norb, nelec, n_steps=1, order=0,
                                       z_representation=False):
    """python/ffsim/trotter/double_factorized.py:25-127."""
    vec = np.asarray(vec, dtype=np.complex128).copy()
    if n_steps == 0:
        return vec
    step_time = time / n_steps
    current = np.eye(norb, dtype=complex)
    for _ in range(n_steps):
        for term, t in trotter_step_iterator(1 + len(diag_coulomb_mats), step_time, order):
            if term == 0:
                current = scipy.linalg.expm(-1j * t * one_body_tensor) @ current
            else:
                U = orbital_rotations[term - 1]
                vec = apply_orbital_rotation(vec, U.T.conj() @ current, norb, nelec, copy=False)
                vec = apply_diag_coulomb_evolution(vec, diag_coulomb_mats[term - 1], t, norb, nelec,
                                                   z_representation=z_representation, copy=False)
                current = U
    vec = apply_orbital_rotation(vec, current, norb, nelec, copy=False)
    if constant:
        vec = vec * np.exp(-1j * time * constant)
    return vec


def simulate_trotter_diag_coulomb_split_op(vec, one_body_tensor, diag_coulomb_mats, constant,
                                           time, norb, nelec, n_steps=1, order=0):
    """python/ffsim/trotter/diagonal_coulomb_split_op.py:25-127: the one-body
    term is merged into the running basis, applied before each diagonal Coulomb
    term (J_aa, J_ab, J_aa) and reset to the identity."""
    vec = np.asarray(vec, dtype=np.complex128).copy()
    if n_steps == 0:
        return vec
    aa, ab = diag_coulomb_mats
    step_time = time / n_steps
    current = np.eye(norb, dtype=complex)
    for _ in range(n_steps):
        for term, t in trotter_step_iterator(2, step_time, order):
            if term == 0:
                current = scipy.linalg.expm(-1j * t * one_body_tensor) @ current
            else:
                vec = apply_orbital_rotation(vec, current, norb, nelec, copy=False)
                vec = apply_diag_coulomb_evolution(vec, (aa, ab, aa), t, norb, nelec, copy=False)
                current = np.eye(norb, dtype=complex)
    vec = apply_orbital_rotation(vec, current, norb, nelec, copy=False)
    if constant:
        vec = vec * np.exp(-1j * time * constant)
    return vec


# ---------------------------------------------------------------------------
# Layer 2: independent closed forms
# ---------------------------------------------------------------------------


def colex_strings(norb: int, nocc: int) -> list[int]:
    """Ascending bitmasks with nocc set bits (closed form of the address order)."""
    return [s for s in range(1 << norb) if bin(s).count("1") == nocc]


def occupations(strings, norb) -> np.ndarray:
    return np.array([[(s >> o) & 1 for o in range(norb)] for s in strings], dtype=float).reshape(
        len(strings), norb
    )


def compound_matrix(U: np.ndarray, nocc: int) -> np.ndarray:
    """Lambda^k(U)[T, S] = det(U[T, S]) over ascending k-subsets (brute force minors)."""
    norb = U.shape[0]
    subsets = [tuple(o for o in range(norb) if (s >> o) & 1) for s in colex_strings(norb, nocc)]
    dim = len(subsets)
    out = np.zeros((dim, dim), dtype=complex)
    for a, T in enumerate(subsets):
        for b, S in enumerate(subsets):
            out[a, b] = np.linalg.det(U[np.ix_(T, S)]) if nocc else 1.0
    return out


def orbital_rotation_closed_form(vec, Ua, Ub, norb, nelec):
    """psi'[a, b] = sum det(Ua[T_a, S_a]) det(Ub[T_b, S_b]) psi[S_a, S_b]."""
    na, nb = nelec
    dim_a, dim_b = math.comb(norb, na), math.comb(norb, nb)
    v = np.asarray(vec).reshape(dim_a, dim_b)
    if Ua is not None:
        v = compound_matrix(Ua, na) @ v
    if Ub is not None:
        v = v @ compound_matrix(Ub, nb).T
    return v.reshape(-1)


def _apply_one_body_term(strings, norb, p, q):
    """a^dag_p a_q on single-spin strings: dense matrix with Jordan-Wigner signs."""
    index = {s: k for k, s in enumerate(strings)}
    m = np.zeros((len(strings), len(strings)))
    for k, s in enumerate(strings):
        if not (s >> q) & 1:
            continue
        sign = (-1) ** bin(s & ((1 << q) - 1)).count("1")
        t = s & ~(1 << q)
        if (t >> p) & 1:
            continue
        sign *= (-1) ** bin(t & ((1 << p) - 1)).count("1")
        t |= 1 << p
        m[index[t], k] += sign
    return m


def one_body_dense(h: np.ndarray, norb: int, nelec) -> np.ndarray:
    """Dense matrix of sum_{pq, sigma} h_pq a^dag_{p sigma} a_{q sigma} on the (na, nb) sector."""
    na, nb = nelec
    sa, sb = colex_strings(norb, na), colex_strings(norb, nb)
    Ha = np.zeros((len(sa), len(sa)), dtype=complex)
    Hb = np.zeros((len(sb), len(sb)), dtype=complex)
    for p in range(norb):
        for q in range(norb):
            if h[p, q] != 0:
                Ha += h[p, q] * _apply_one_body_term(sa, norb, p, q)
                Hb += h[p, q] * _apply_one_body_term(sb, norb, p, q)
    return np.kron(Ha, np.eye(len(sb))) + np.kron(np.eye(len(sa)), Hb)


def molecular_hamiltonian_dense(h1, h2, constant, norb, nelec) -> np.ndarray:
    """Dense sum_pq h1_pq E_pq + 1/2 sum (pq|rs) (E_pq E_rs - delta_qr E_ps) +
    constant on the (na, nb) sector, with E_pq the spin-summed one-body
    excitation from one_body_dense (chemists' order for h2)."""
    def e(p, q):
        m = np.zeros((norb, norb))
        m[p, q] = 1
        return one_body_dense(m, norb, nelec)

    E = [[e(p, q) for q in range(norb)] for p in range(norb)]
    H = constant * np.eye(E[0][0].shape[0], dtype=complex)
    for p in range(norb):
        for q in range(norb):
            H += h1[p, q] * E[p][q]
            for r in range(norb):
                for s in range(norb):
                    if h2[p, q, r, s]:
                        H += 0.5 * h2[p, q, r, s] * (E[p][q] @ E[r][s] - (q == r) * E[p][s])
    return H


def diag_coulomb_energies(mat, norb, nelec, z_representation=False, rows=None) -> np.ndarray:
    """E(a, b) on every determinant from occupation numbers (closed form).

    Number representation: sum_{ij, sigma tau} J^{sigma tau}_ij n_i n_j / 2 with
    J given as one matrix or (aa, ab, bb) (None -> 0).  Z representation:
    (1/8) sum^*_{ij sigma tau} J z_i z_j with z = 1 - 2 n, the starred sum
    excluding i == j within one spin.

    With ``rows`` (a slice of alpha addresses) only that block is formed and it
    is returned as a 2-D (alpha rows x dim_b) array; otherwise the whole state's
    energies, flattened.
    """
    na, nb = nelec
    occ_a = occupations(colex_strings(norb, na), norb)
    occ_b = occupations(colex_strings(norb, nb), norb)
    if rows is not None:
        occ_a = occ_a[rows]
    if isinstance(mat, np.ndarray):
        aa = ab = bb = mat
        ba = mat
    else:
        aa, ab, bb = (np.zeros((norb, norb)) if m is None else np.asarray(m, float) for m in mat)
        ba = ab.T
    if not z_representation:
        ea = 0.5 * np.einsum("ai,ij,aj->a", occ_a, aa, occ_a)
        eb = 0.5 * np.einsum("bi,ij,bj->b", occ_b, bb, occ_b)
        cross = 0.5 * (occ_a @ ab @ occ_b.T + (occ_b @ ba @ occ_a.T).T)
        e = ea[:, None] + eb[None, :] + cross
        return e if rows is not None else e.reshape(-1)
    za, zb = 1 - 2 * occ_a, 1 - 2 * occ_b
    off_aa = aa - np.diag(np.diag(aa))
    off_bb = bb - np.diag(np.diag(bb))
    ea = np.einsum("ai,ij,aj->a", za, off_aa, za)
    eb = np.einsum("bi,ij,bj->b", zb, off_bb, zb)
    cross = za @ ab @ zb.T + (zb @ ba @ za.T).T
    e = 0.125 * (ea[:, None] + eb[None, :] + cross)
    return e if rows is not None else e.reshape(-1)


def num_op_energies(coeffs_a, coeffs_b, norb, nelec) -> np.ndarray:
    na, nb = nelec
    occ_a = occupations(colex_strings(norb, na), norb)
    occ_b = occupations(colex_strings(norb, nb), norb)
    ea = occ_a @ coeffs_a if coeffs_a is not None else np.zeros(len(occ_a))
    eb = occ_b @ coeffs_b if coeffs_b is not None else np.zeros(len(occ_b))
    return (ea[:, None] + eb[None, :]).reshape(-1)


def random_unitary(dim: int, rng) -> np.ndarray:
    """Haar unitary (python/ffsim/random/random.py:84-106 convention)."""
    z = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))


def random_state_vector(dim: int, rng) -> np.ndarray:
    """python/ffsim/random/random.py:24-47."""
    vec = rng.standard_normal(dim).astype(complex)
    vec += 1j * rng.standard_normal(dim)
    vec /= np.linalg.norm(vec)
    return vec


def random_symmetric_uniform(dim: int, rng, scale: float = 2.0) -> np.ndarray:
    """python/ffsim/random/random.py:535-546 (mean 0)."""
    vals = rng.uniform(-0.5 * scale, 0.5 * scale, size=dim * (dim + 1) // 2)
    mat = np.zeros((dim, dim))
    rows, cols = np.triu_indices(dim)
    mat[rows, cols] = vals
    mat[cols, rows] = vals
    return mat
