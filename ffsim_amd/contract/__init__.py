"""Contraction drop-ins (python/ffsim/contract/)."""

from ffsim_amd.contract.diag_coulomb import contract_diag_coulomb, diag_coulomb_linop
from ffsim_amd.contract.num_op_sum import contract_num_op_sum, num_op_sum_linop

__all__ = [
    "contract_diag_coulomb",
    "contract_num_op_sum",
    "diag_coulomb_linop",
    "num_op_sum_linop",
]
