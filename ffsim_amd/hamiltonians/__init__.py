"""Hamiltonians with GPU linear operators (python/ffsim/hamiltonians/)."""

from ffsim_amd.hamiltonians.diagonal_coulomb_hamiltonian import DiagonalCoulombHamiltonian
from ffsim_amd.hamiltonians.double_factorized_hamiltonian import DoubleFactorizedHamiltonian

__all__ = ["DiagonalCoulombHamiltonian", "DoubleFactorizedHamiltonian"]
