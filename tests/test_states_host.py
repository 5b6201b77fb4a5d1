"""Address <-> bitstring conversion (python/ffsim/states/bitstring.py), pinned
by the reference's known answers (tests/python/states/bitstring_test.py:24-97,
:282-299, held in tests/golden/bitstrings.json), by the docstring example of
strings_to_addresses (bitstring.py:270-318) and exhaustively by the oracle's
string tables for small norb."""

from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

from ffsim_amd import states as S
from oracle import oracle as O

KNOWN = json.loads((Path(__file__).resolve().parent / "golden" / "bitstrings.json").read_text())
B = S.BitstringType


def test_known_answers():
    assert S.addresses_to_strings(range(3), 3, 2) == KNOWN["spinless_norb3_nelec2"]
    assert S.addresses_to_strings(range(9), 3, (2, 1)) == KNOWN["spinful_norb3_nelec21_concat"]
    k = KNOWN["large_address"]
    got = S.addresses_to_strings(k["addresses"], k["norb"], tuple(k["nelec"]),
                                 bitstring_type=B.STRING)
    assert got == k["strings"]
    np.testing.assert_array_equal(
        S.strings_to_addresses(k["strings"], k["norb"], tuple(k["nelec"])), k["addresses"])


def test_docstring_example():
    ints = [0b001011, 0b010101, 0b100101, 0b010110, 0b100110]
    strs = ["001011", "010101", "100101", "010110", "100110"]
    bits = S.convert_bitstring_type(ints, B.INT, B.BIT_ARRAY, 6)
    for x in (ints, strs, bits):
        np.testing.assert_array_equal(S.strings_to_addresses(x, 3, (2, 1)), [0, 4, 5, 7, 8])


@pytest.mark.parametrize("norb", range(1, 8))
def test_exhaustive_against_oracle_tables(norb):
    for k in range(norb + 1):
        strings = [int(s) for s in O.cached_strings(norb, k)]
        assert S.addresses_to_strings(range(len(strings)), norb, k) == strings
        np.testing.assert_array_equal(S.strings_to_addresses(strings, norb, k),
                                      np.arange(len(strings)))
    na, nb = norb // 2, (norb + 1) // 2
    dim = math.comb(norb, na) * math.comb(norb, nb)
    a, b = S.addresses_to_strings(range(dim), norb, (na, nb), concatenate=False)
    sa, sb = O.cached_strings(norb, na), O.cached_strings(norb, nb)
    assert a == [int(sa[i // len(sb)]) for i in range(dim)]
    assert b == [int(sb[i % len(sb)]) for i in range(dim)]


def test_formats_round_trip():
    ints = [0, 5, 6, 9]
    for fmt in B:
        x = S.convert_bitstring_type(ints, B.INT, fmt, 4)
        assert S.convert_bitstring_type(x, fmt, B.INT, 4) == ints
    assert S.convert_bitstring_type(["0101"], B.STRING, B.BIT_ARRAY, 4).tolist() == [
        [False, True, False, True]]
    assert S.restrict_bitstrings([0b1101], [0, 2, 3], B.INT) == [0b111]
    assert S.restrict_bitstrings(["1101"], [0, 2, 3], B.STRING) == ["111"]
    assert S.bitstring_to_occupied_orbitals(0b1101) == [0, 2, 3]
    with pytest.raises(TypeError):
        S.convert_bitstring_type(ints, "int", B.STRING, 4)
    with pytest.raises(ValueError):
        S.strings_to_addresses([0b111], 3, 2)


def test_wide_strings_round_trip():
    """More than 62 orbitals take the Python-integer path; both paths agree
    where they overlap."""
    for norb, nocc in ((64, 2), (70, 3), (40, 3)):
        n = math.comb(norb, nocc)
        addrs = np.unique(np.linspace(0, n - 1, 97).astype(np.int64))
        strings = S.addresses_to_strings(addrs, norb, nocc)
        assert all(bin(x).count("1") == nocc and x >> norb == 0 for x in strings)
        assert strings == sorted(strings)
        np.testing.assert_array_equal(S.strings_to_addresses(strings, norb, nocc), addrs)
