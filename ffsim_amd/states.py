"""State dimensions and reference states (python/ffsim/states/)."""

from __future__ import annotations

import math
from enum import Enum, auto

import numpy as np


def dims(norb: int, nelec: tuple[int, int]) -> tuple[int, int]:
    """(dim_a, dim_b) (python/ffsim/states/dimensions.py:18-33)."""
    n_alpha, n_beta = nelec
    return math.comb(norb, n_alpha), math.comb(norb, n_beta)


def dim(norb: int, nelec) -> int:
    """Dimension of the FCI space (python/ffsim/states/dimensions.py:36-50)."""
    if isinstance(nelec, int):
        return math.comb(norb, nelec)
    n_alpha, n_beta = nelec
    return math.comb(norb, n_alpha) * math.comb(norb, n_beta)


def hartree_fock_state(norb: int, nelec) -> np.ndarray:
    """One-hot state at address 0, i.e. the lowest orbitals occupied
    (python/ffsim/states/slater.py:122-138)."""
    vec = np.zeros(dim(norb, nelec), dtype=complex)
    vec[0] = 1
    return vec


# ---------------------------------------------------------------------------
# Address <-> bitstring conversion (python/ffsim/states/bitstring.py).  The
# address of a string with occupied orbitals c_1 < ... < c_k is its colex rank
# sum_t C(c_t, t); a spinful address is a * dim_b + b and its bitstring puts
# the beta string above the alpha one, (beta << norb) | alpha.


class BitstringType(Enum):
    """Bitstring formats (bitstring.py:21-45): STRING ("0101", most significant
    bit first), INT (5) and BIT_ARRAY (a boolean row, most significant first)."""

    STRING = auto()
    INT = auto()
    BIT_ARRAY = auto()


def convert_bitstring_type(strings, input_type: BitstringType, output_type: BitstringType,
                           length: int):
    """Convert bitstrings between formats (bitstring.py:48-114)."""
    for t, name in ((input_type, "input_type"), (output_type, "output_type")):
        if not isinstance(t, BitstringType):
            raise TypeError(
                f"{name}: expected a BitstringType (STRING, INT or BIT_ARRAY), got {t}.")
    if input_type is output_type:
        return strings
    if input_type is BitstringType.BIT_ARRAY:
        ints = [sum(1 << i for i, bit in enumerate(row[::-1]) if bit) for row in strings]
    elif input_type is BitstringType.STRING:
        ints = [int(s, 2) for s in strings]
    else:
        ints = [int(s) for s in strings]
    if output_type is BitstringType.INT:
        return ints
    if output_type is BitstringType.STRING:
        return [format(s, f"0{length}b") for s in ints]
    if not ints:
        return np.empty((0, length), dtype=bool)
    return np.array([[(s >> i) & 1 for i in range(length - 1, -1, -1)] for s in ints],
                    dtype=bool)


def concatenate_bitstrings(strings_a, strings_b, bitstring_type: BitstringType, length: int):
    """Beta strings above alpha strings (bitstring.py:145-170)."""
    if bitstring_type is BitstringType.STRING:
        return [b + a for a, b in zip(strings_a, strings_b)]
    if bitstring_type is BitstringType.INT:
        return [(b << length) | a for a, b in zip(strings_a, strings_b)]
    return np.concatenate([strings_b, strings_a], axis=1)


def restrict_bitstrings(strings, indices, bitstring_type: BitstringType):
    """Keep the bits at ``indices`` (bitstring.py:117-142)."""
    if bitstring_type is BitstringType.STRING:
        return ["".join(s[-1 - i] for i in indices[::-1]) for s in strings]
    if bitstring_type is BitstringType.INT:
        return [sum(((s >> idx) & 1) << i for i, idx in enumerate(indices)) for s in strings]
    return np.asarray(strings)[:, [-1 - i for i in indices[::-1]]]


def _binom_table(norb: int) -> np.ndarray:
    """C(c, t) for 0 <= c < norb and 0 <= t <= norb + 1, as int64 (norb <= 62)."""
    table = np.zeros((max(norb, 1), norb + 2), dtype=np.int64)
    for c in range(norb):
        for t in range(c + 1):
            table[c, t] = math.comb(c, t)
    return table


def _colex_unrank(addresses, norb: int, nocc: int) -> list[int]:
    """Strings of the given colex ranks: for t = nocc..1 the largest c with
    C(c, t) <= the remaining rank is the t-th occupied orbital."""
    addr = np.asarray(addresses, dtype=np.int64).reshape(-1)
    if addr.size and (addr.min() < 0 or addr.max() >= math.comb(norb, nocc)):
        raise ValueError(f"address out of range for norb={norb}, nocc={nocc}")
    if norb > 62:
        out = []
        for a in addr:
            s, r, c = 0, int(a), norb
            for t in range(nocc, 0, -1):
                c -= 1
                while math.comb(c, t) > r:
                    c -= 1
                s |= 1 << c
                r -= math.comb(c, t)
            out.append(s)
        return out
    table = _binom_table(norb)
    s = np.zeros(addr.shape, dtype=np.int64)
    r = addr.copy()
    for t in range(nocc, 0, -1):
        c = np.searchsorted(table[:, t], r, side="right") - 1
        s |= np.left_shift(np.int64(1), c)
        r -= table[c, t]
    return [int(x) for x in s]


def _colex_rank(strings, norb: int, nocc: int) -> np.ndarray:
    """Colex ranks sum_t C(c_t, t) of strings with nocc of norb bits set."""
    ints = [int(x) for x in strings]
    if any(x < 0 or x >> norb or bin(x).count("1") != nocc for x in ints):
        raise ValueError(f"a bitstring is not a string of {nocc} in {norb} orbitals")
    if norb > 62:
        out = np.zeros(len(ints), dtype=np.int64)
        for k, x in enumerate(ints):
            rank, t = 0, 0
            while x:
                t += 1
                rank += math.comb((x & -x).bit_length() - 1, t)
                x &= x - 1
            out[k] = rank
        return out
    table = _binom_table(norb)
    s = np.asarray(ints, dtype=np.int64)
    rank = np.zeros(s.shape, dtype=np.int64)
    t = np.zeros(s.shape, dtype=np.int64)
    for c in range(norb):
        bit = (s >> c) & 1
        t += bit
        rank += bit * table[c, t]
    return rank


def addresses_to_strings(addresses, norb: int, nelec, concatenate: bool = True,
                         bitstring_type: BitstringType = BitstringType.INT):
    """Bitstrings of state-vector addresses (bitstring.py:173-263).  Spinful:
    concatenated (beta above alpha) or, with ``concatenate=False``, the pair
    (alpha strings, beta strings)."""
    addresses = np.asarray(addresses, dtype=np.int64).reshape(-1)
    if isinstance(nelec, int):
        return convert_bitstring_type(_colex_unrank(addresses, norb, nelec), BitstringType.INT,
                                      bitstring_type, norb)
    n_alpha, n_beta = nelec
    addr_a, addr_b = np.divmod(addresses, math.comb(norb, n_beta))
    strings_a = convert_bitstring_type(_colex_unrank(addr_a, norb, n_alpha), BitstringType.INT,
                                       bitstring_type, norb)
    strings_b = convert_bitstring_type(_colex_unrank(addr_b, norb, n_beta), BitstringType.INT,
                                       bitstring_type, norb)
    if concatenate:
        return concatenate_bitstrings(strings_a, strings_b, bitstring_type, norb)
    return strings_a, strings_b


def strings_to_addresses(strings, norb: int, nelec) -> np.ndarray:
    """State-vector addresses of bitstrings (bitstring.py:266-361); the
    format is read from the input (2-D array: bits; str: strings; else ints)."""
    if not len(strings):
        return np.array([])
    if isinstance(strings, np.ndarray) and strings.ndim == 2:
        kind = BitstringType.BIT_ARRAY
    elif isinstance(strings[0], str):
        kind = BitstringType.STRING
    else:
        kind = BitstringType.INT
    ints = convert_bitstring_type(strings, kind, BitstringType.INT, norb)
    if isinstance(nelec, int):
        return _colex_rank(ints, norb, nelec)
    n_alpha, n_beta = nelec
    mask = (1 << norb) - 1
    addr_a = _colex_rank([s & mask for s in ints], norb, n_alpha)
    addr_b = _colex_rank([s >> norb for s in ints], norb, n_beta)
    return addr_a * math.comb(norb, n_beta) + addr_b


def bitstring_to_occupied_orbitals(bitstring: int) -> list[int]:
    """Positions of the set bits (bitstring.py:364-366)."""
    return [i for i in range(int(bitstring).bit_length()) if (int(bitstring) >> i) & 1]
