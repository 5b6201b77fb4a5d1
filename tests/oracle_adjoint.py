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


def five_point_derivative(func, theta, direction, h=1e-3):
    """d/de func(theta + e * direction) at e = 0 by the five-point central stencil,
    (-f(2h) + 8 f(h) - 8 f(-h) + f(-2h)) / (12 h), truncation O(h^4)."""
    theta = np.asarray(theta, dtype=float)
    d = np.asarray(direction, dtype=float)
    f = lambda s: func(theta + s * d)  # noqa: E731
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def five_point_gradient(func, theta, h=1e-3, indices=None):
    """Five-point central finite differences along the coordinate axes."""
    theta = np.asarray(theta, dtype=float)
    indices = range(len(theta)) if indices is None else indices
    out = []
    for i in indices:
        e = np.zeros_like(theta)
        e[i] = 1.0
        out.append(five_point_derivative(func, theta, e, h))
    return np.array(out)
