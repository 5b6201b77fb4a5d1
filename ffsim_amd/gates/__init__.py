"""Gate drop-ins (python/ffsim/gates/)."""

from ffsim_amd.gates.diag_coulomb import apply_diag_coulomb_evolution
from ffsim_amd.gates.num_op_sum import apply_num_op_sum_evolution
from ffsim_amd.gates.orbital_rotation import apply_orbital_rotation

__all__ = [
    "apply_diag_coulomb_evolution",
    "apply_num_op_sum_evolution",
    "apply_orbital_rotation",
]
