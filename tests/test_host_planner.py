"""Host planner, fiber and group tables, checked on the CPU.

tests/native/pass_emulator.cpp runs the exact programs the library builds
for the GPU fiber kernel (window passes with transposed stores, register
sub-passes, layer pairs, group tables) with plain loops; the result must
match the oracle.  Small windows force many passes and sub-windows.
"""

from __future__ import annotations

import ctypes
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def emulator(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("emu") / "libpass_emulator.so"
    subprocess.run(
        ["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
         str(ROOT / "tests/native/pass_emulator.cpp"),
         str(ROOT / "ffsim_amd/csrc/host_core.cpp"), "-o", str(out)],
        check=True, capture_output=True)
    lib = ctypes.CDLL(str(out))
    P = ctypes.c_void_p
    lib.emulate.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int,
                            ctypes.c_int]
    lib.emulate_order.argtypes = lib.emulate.argtypes + [ctypes.c_int]
    lib.decompose.argtypes = [P, ctypes.c_int, ctypes.c_int, P, P, P, P, P]
    lib.decompose.restype = ctypes.c_int
    lib.plan_counts.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P]
    lib.emulate_one_body.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                     ctypes.c_int]
    lib.emulate_one_body.restype = ctypes.c_int
    return lib


def _p(a):
    return np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("merge", [0, 1])
@pytest.mark.parametrize("norb,nelec,w", [
    (4, (2, 2), 4), (5, (2, 3), 5), (6, (3, 2), 4), (7, (0, 3), 4), (7, (7, 2), 4),
    (8, (4, 4), 5), (8, (3, 5), 6), (9, (2, 7), 5), (10, (5, 4), 6), (12, (6, 6), 8),
    (12, (6, 6), 12), (13, (6, 5), 12)])
def test_emulated_passes_match_oracle(emulator, norb, nelec, w, merge, rng):
    """merge=1: groups with more than half the sub-window occupied are stored as
    complemented groups of the mirrored class (the c64 tables)."""
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    vec = O.random_state_vector(da * db, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    expected = O.apply_orbital_rotation(vec, (Ua, Ub), norb, nelec)
    x = vec.copy()
    work = np.zeros_like(x)
    p = lambda a: np.ascontiguousarray(a).ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    ua, ub = np.ascontiguousarray(Ua), np.ascontiguousarray(Ub)
    emulator.emulate(p(x), p(work), norb, nelec[0], nelec[1], p(ua), p(ub), w, merge)
    np.testing.assert_allclose(x, expected, atol=1e-12, rtol=0)


@pytest.mark.parametrize("norb,nelec,w", [(8, (4, 4), 5), (10, (5, 4), 6), (13, (6, 5), 12)])
def test_emulated_passes_in_clements_order(emulator, norb, nelec, w, rng):
    """FFSIM_AMD_GIVENS_ORDER=clements: ffsim's rotation order through the same passes."""
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    vec = O.random_state_vector(da * db, rng)
    Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
    expected = O.apply_orbital_rotation(vec, (Ua, Ub), norb, nelec)
    x, work = vec.copy(), np.zeros_like(vec)
    emulator.emulate_order(_p(x), _p(work), norb, nelec[0], nelec[1], _p(Ua), _p(Ub), w, 0, 1)
    np.testing.assert_allclose(x, expected, atol=1e-12, rtol=0)


def _decompose(emulator, U, clements):
    n = U.shape[0]
    m = n * (n - 1) // 2
    c, s = np.zeros(m), np.zeros(2 * m)
    i, j = np.zeros(m, dtype=np.int32), np.zeros(m, dtype=np.int32)
    d = np.zeros(2 * n)
    k = emulator.decompose(_p(U), n, clements, _p(c), _p(s), _p(i), _p(j), _p(d))
    return c[:k], (s[0::2] + 1j * s[1::2])[:k], i[:k], j[:k], d[0::2] + 1j * d[1::2]


@pytest.mark.parametrize("n", [2, 3, 5, 8, 16, 18])
def test_triangle_decomposition_reconstructs(emulator, n, rng):
    """U = D G_N^* ... G_1^*, G = [[c, s], [-conj(s), c]] on adjacent (i, j):
    the same form as ffsim.linalg.givens_decomposition, rotations in Reck's
    order (each row nulled left to right, rows from the bottom)."""
    U = O.random_unitary(n, rng)
    c, s, i, j, d = _decompose(emulator, U, 0)
    assert len(c) == n * (n - 1) // 2
    assert np.all(np.abs(i - j) == 1)
    np.testing.assert_allclose(np.abs(d), 1.0, atol=1e-12)
    acc = np.diag(d)
    for ck, sk, ik, jk in zip(c[::-1], s[::-1], i[::-1], j[::-1]):
        G = np.eye(n, dtype=complex)
        G[ik, ik], G[ik, jk], G[jk, ik], G[jk, jk] = ck, sk, -np.conj(sk), ck
        acc = acc @ G.conj()
    np.testing.assert_allclose(acc, U, atol=1e-12)
    # the first sweep walks row n-1 from orbital pair (0, 1) to (n-2, n-1)
    np.testing.assert_array_equal(np.minimum(i, j)[: n - 1], np.arange(n - 1))


def test_triangle_order_needs_fewer_passes(emulator, rng):
    """At 12-orbital windows the triangle order needs 3 passes at norb 17 and 18
    (ffsim's order 4) and no more register sub-passes anywhere."""
    for n in range(13, 21):
        U = O.random_unitary(n, rng)
        got = {}
        for clements in (0, 1):
            np_, ns = ctypes.c_int(), ctypes.c_int()
            emulator.plan_counts(_p(U), n, 12, clements, ctypes.byref(np_), ctypes.byref(ns))
            got[clements] = (np_.value, ns.value)
        assert got[0][0] <= got[1][0] and got[0][1] <= got[1][1], (n, got)
        if n in (17, 18):
            assert got[0][0] == 3 and got[1][0] == 4, (n, got)


@pytest.mark.parametrize("norb,nelec,w,n_passes", [
    (4, (2, 1), 4, 1), (6, (3, 2), 4, 3), (7, (3, 4), 5, 3), (8, (4, 4), 6, 3),
    (9, (0, 5), 6, 3), (10, (5, 3), 8, 3), (8, (8, 2), 6, 3), (12, (6, 6), 12, 1)])
def test_emulated_one_body_window_passes_match_oracle(emulator, norb, nelec, w, n_passes, rng):
    """The one-body window passes (windows A+B, B+C and the split window A+C
    whose fibers carry the sign of their electrons on B) against the oracle's
    restatement of pyscf's contract_1e over both spin sectors."""
    da, db = math.comb(norb, nelec[0]), math.comb(norb, nelec[1])
    vec = O.random_state_vector(da * db, rng)
    M = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
    out = np.zeros_like(vec)
    got = emulator.emulate_one_body(_p(vec), _p(out), norb, nelec[0], nelec[1], _p(M), w)
    assert got == n_passes
    np.testing.assert_allclose(out, O.contract_one_body(vec, M, norb, nelec), atol=1e-12, rtol=0)


def test_one_body_passes_need_a_wide_enough_window(emulator):
    vec = np.zeros(math.comb(10, 5) ** 2, dtype=complex)
    M = np.eye(10, dtype=complex)
    assert emulator.emulate_one_body(_p(vec), _p(vec.copy()), 10, 5, 5, _p(M), 6) == 0


def test_emulated_passes_random_shapes(emulator):
    """Random shapes, windows and rotation orders (hypothesis): the emulated
    fused passes and one-body window passes against the oracle."""
    hyp = pytest.importorskip("hypothesis")
    st = hyp.strategies

    @hyp.settings(max_examples=40, deadline=None, derandomize=True)
    @hyp.given(norb=st.integers(2, 10), data=st.data())
    def check(norb, data):
        na = data.draw(st.integers(0, norb))
        nb = data.draw(st.integers(0, norb))
        w = data.draw(st.integers(2, norb))
        clements = data.draw(st.integers(0, 1))
        merge = data.draw(st.integers(0, 1))
        seed = data.draw(st.integers(0, 2**31 - 1))
        rng = np.random.default_rng(seed)
        da, db = math.comb(norb, na), math.comb(norb, nb)
        vec = O.random_state_vector(da * db, rng)
        Ua, Ub = O.random_unitary(norb, rng), O.random_unitary(norb, rng)
        x, work = vec.copy(), np.zeros_like(vec)
        emulator.emulate_order(_p(x), _p(work), norb, na, nb, _p(Ua), _p(Ub), w, merge, clements)
        np.testing.assert_allclose(x, O.apply_orbital_rotation(vec, (Ua, Ub), norb, (na, nb)),
                                   atol=1e-12, rtol=0)
        M = rng.standard_normal((norb, norb)) + 1j * rng.standard_normal((norb, norb))
        out = np.zeros_like(vec)
        n = emulator.emulate_one_body(_p(vec), _p(out), norb, na, nb, _p(M), w)
        assert n == (1 if norb <= w else 3 if 2 * (norb - w) <= w else 0)
        if n:
            np.testing.assert_allclose(out, O.contract_one_body(vec, M, norb, (na, nb)),
                                       atol=1e-12, rtol=0)

    check()
