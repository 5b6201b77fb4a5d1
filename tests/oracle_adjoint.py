"""The adjoint sweep's primitive steps computed by the CPU oracle (test helper)."""

from __future__ import annotations

import numpy as np

from oracle import oracle as O


class OracleAdjointOps:
    """The sweep's primitive steps on numpy vectors, computed by the oracle."""

    def __init__(self, norb, nelec):
        self.norb, self.nelec = norb, tuple(nelec)
        self.occ_a = O.occupations(O.colex_strings(norb, nelec[0]), norb)
        self.occ_b = O.occupations(O.colex_strings(norb, nelec[1]), norb)

    def clone(self, t):
        return np.array(t, dtype=complex)

    def rotate(self, t, U):
        return O.apply_orbital_rotation(t, U, self.norb, self.nelec, copy=False)

    def diag(self, t, j_aa, j_ab, time):
        return O.apply_diag_coulomb_evolution(t, (j_aa, j_ab, j_aa), time, self.norb, self.nelec,
                                              copy=False)

    def transition_rdm(self, lam, psi):
        return O.transition_rdm(lam, psi, self.norb, self.nelec)

    def phase_reductions(self, psi, lam):
        return O.diag_coulomb_phase_reductions(psi, lam, self.norb, self.nelec)

    def host_rdm(self, x):
        return x

    def host_phase(self, x):
        return x
