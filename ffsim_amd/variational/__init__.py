"""Variational ansatz operators (python/ffsim/variational/)."""

from ffsim_amd.variational.givens import GivensAnsatzOp
from ffsim_amd.variational.num_num import NumNumAnsatzOpSpinBalanced
from ffsim_amd.variational.ucj_angles_spin_balanced import UCJAnglesOpSpinBalanced
from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced
from ffsim_amd.variational.ucj_spin_unbalanced import UCJOpSpinUnbalanced
from ffsim_amd.variational.ucj_spinless import UCJOpSpinless

__all__ = [
    "GivensAnsatzOp",
    "NumNumAnsatzOpSpinBalanced",
    "UCJAnglesOpSpinBalanced",
    "UCJOpSpinBalanced",
    "UCJOpSpinUnbalanced",
    "UCJOpSpinless",
]
