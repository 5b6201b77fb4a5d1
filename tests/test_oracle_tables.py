"""Pin the integer tables: determinant strings, occupation lists, pair tables.

Bit-exact checks against the reference's known answers
(tests/python/states/bitstring_test.py), against closed forms, and between
the oracle's pyscf restatement and the library's host tables.
"""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden"


def _known():
    return json.loads((GOLDEN / "bitstrings.json").read_text())


def test_spinless_known_answer():
    # bitstring_test.py:24-49: norb=3, nelec=2 -> 0b011, 0b101, 0b110
    assert [int(s) for s in O.make_strings(range(3), 2)] == _known()["spinless_norb3_nelec2"]


def test_spinful_known_answer():
    # bitstring_test.py:52-97: address a*dim_b + b -> (beta << norb) | alpha
    norb, (na, nb) = 3, (2, 1)
    sa, sb = O.make_strings(range(norb), na), O.make_strings(range(norb), nb)
    concat = [int(b) << norb | int(a) for a in sa for b in sb]
    assert concat == _known()["spinful_norb3_nelec21_concat"]


def test_large_address_known_answer():
    # bitstring_test.py:282-299: norb=33, nelec=(3,3), addresses 29767920..29767929
    k = _known()["large_address"]
    norb, (na, nb) = k["norb"], k["nelec"]
    dim_b = math.comb(norb, nb)

    def unrank(addr, n, kocc):
        # colex unranking: largest c with C(c, t) <= remaining, for t = kocc..1
        s, r = 0, addr
        for t in range(kocc, 0, -1):
            c = t - 1
            while math.comb(c + 1, t) <= r:
                c += 1
            s |= 1 << c
            r -= math.comb(c, t)
        return s

    sa = O.make_strings(range(norb), na)
    sb = O.make_strings(range(norb), nb)
    for addr, expected in zip(k["addresses"], k["strings"]):
        a, b = divmod(addr, dim_b)
        assert int(sb[b]) << norb | int(sa[a]) == int(expected, 2)
        assert unrank(b, norb, nb) << norb | unrank(a, norb, na) == int(expected, 2)


@pytest.mark.parametrize("norb", range(0, 10))
def test_strings_are_ascending_k_subsets(norb):
    for k in range(norb + 1):
        strings = [int(s) for s in O.make_strings(range(norb), k)]
        assert strings == O.colex_strings(norb, k)


def test_occslst_matches_strings():
    for norb in range(1, 9):
        for k in range(norb + 1):
            occ = O.cached_occslst(norb, k)
            strings = O.cached_strings(norb, k)
            for s, row in zip(strings, occ):
                assert [o for o in range(norb) if (int(s) >> o) & 1] == [int(x) for x in row]


@pytest.mark.parametrize("norb", range(2, 10))
def test_zero_one_indices_closed_form(norb):
    """slice1 = ascending addresses with i occupied / j empty; slice2 = partners."""
    for k in range(norb + 1):
        strings = O.colex_strings(norb, k)
        index = {s: a for a, s in enumerate(strings)}
        for i in range(norb - 1):
            for pair in ((i, i + 1), (i + 1, i)):
                a, b = pair
                first = [index[s] for s in strings if (s >> a) & 1 and not (s >> b) & 1]
                second = [index[s ^ (1 << a) ^ (1 << b)] for s in strings
                          if (s >> a) & 1 and not (s >> b) & 1]
                got = O.zero_one_subspace_indices(norb, k, pair)
                assert [int(x) for x in got] == first + second


def test_library_tables_bit_exact():
    from ffsim_amd import _cistring

    for norb in range(1, 11):
        for k in range(norb + 1):
            np.testing.assert_array_equal(_cistring.make_strings(range(norb), k),
                                          O.make_strings(range(norb), k))
            np.testing.assert_array_equal(_cistring.gen_occslst(range(norb), k).astype(np.int64),
                                          O.cached_occslst(norb, k).astype(np.int64))
            if norb >= 2:
                for i in range(norb - 1):
                    for pair in ((i, i + 1), (i + 1, i)):
                        np.testing.assert_array_equal(
                            _cistring.zero_one_subspace_indices(norb, k, pair).astype(np.int64),
                            O.zero_one_subspace_indices(norb, k, pair).astype(np.int64),
                        )


def test_golden_tables():
    g = np.load(GOLDEN / "golden.npz")
    for n in range(1, 9):
        for k in range(n + 1):
            np.testing.assert_array_equal(g[f"strings_{n}_{k}"], O.cached_strings(n, k))
    for k in range(1, 8):
        for i in range(7):
            for pair in ((i, i + 1), (i + 1, i)):
                np.testing.assert_array_equal(
                    g[f"zero_one_8_{k}_{pair[0]}_{pair[1]}"],
                    O.zero_one_subspace_indices(8, k, pair).astype(np.int64),
                )
