"""Variational ansatz operators (python/ffsim/variational/)."""

from ffsim_amd.variational.ucj_spin_balanced import UCJOpSpinBalanced

__all__ = ["UCJOpSpinBalanced"]
