"""Variational ansatz operators (python/ffsim/variational/)."""

from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced
from ffsim_amd.variational.ucj_spin_unbalanced import UCJOpSpinUnbalanced
from ffsim_amd.variational.ucj_spinless import UCJOpSpinless

__all__ = ["UCJOpSpinBalanced", "UCJOpSpinUnbalanced", "UCJOpSpinless"]
