"""ffsim_amd: MI355X (gfx950) drop-ins for ffsim's gate-application hot path.

The public names follow ffsim's (qiskit-community/ffsim): gate functions,
protocol dispatch, the spin-balanced UCJ operator and the Hamiltonians whose
linear operators are on the path.  States are numpy arrays (copied to HBM
and back) or torch tensors on the GPU (kept resident).
"""

from ffsim_amd import contract, linalg, random
from ffsim_amd.gates import (
    apply_diag_coulomb_evolution,
    apply_num_op_sum_evolution,
    apply_orbital_rotation,
)
from ffsim_amd.hamiltonians import DiagonalCoulombHamiltonian, DoubleFactorizedHamiltonian
from ffsim_amd.protocols import apply_unitary, approx_eq, diag, linear_operator, trace
from ffsim_amd.states import (
    BitstringType,
    addresses_to_strings,
    dim,
    dims,
    hartree_fock_state,
    strings_to_addresses,
)
from ffsim_amd.trotter import (
    simulate_trotter_diag_coulomb_split_op,
    simulate_trotter_double_factorized,
)
from ffsim_amd.variational import (
    GivensAnsatzOp,
    NumNumAnsatzOpSpinBalanced,
    UCJAnglesOpSpinBalanced,
    UCJOpSpinBalanced,
    UCJOpSpinless,
    UCJOpSpinUnbalanced,
)

__all__ = [
    "BitstringType",
    "addresses_to_strings",
    "strings_to_addresses",
    "DiagonalCoulombHamiltonian",
    "DoubleFactorizedHamiltonian",
    "GivensAnsatzOp",
    "NumNumAnsatzOpSpinBalanced",
    "UCJAnglesOpSpinBalanced",
    "UCJOpSpinBalanced",
    "UCJOpSpinUnbalanced",
    "UCJOpSpinless",
    "apply_diag_coulomb_evolution",
    "apply_num_op_sum_evolution",
    "apply_orbital_rotation",
    "apply_unitary",
    "approx_eq",
    "diag",
    "contract",
    "dim",
    "dims",
    "hartree_fock_state",
    "linalg",
    "linear_operator",
    "random",
    "simulate_trotter_diag_coulomb_split_op",
    "simulate_trotter_double_factorized",
    "trace",
]
