"""Protocol dispatch (python/ffsim/protocols/)."""

from __future__ import annotations

from typing import Any

import numpy as np


def apply_unitary(vec, obj: Any, norb: int, nelec, copy: bool = True):
    """Apply a unitary transformation to a vector.

    Same dispatch as ffsim.apply_unitary
    (python/ffsim/protocols/apply_unitary_protocol.py:48-88): calls
    ``obj._apply_unitary_`` and raises TypeError when the object has none or
    it returns NotImplemented.
    """
    method = getattr(obj, "_apply_unitary_", None)
    if method is not None:
        result = method(vec, norb=norb, nelec=nelec, copy=copy)
        if result is not NotImplemented:
            return result
    raise TypeError(
        f"apply_unitary: {type(obj)} has no _apply_unitary_, or it returned NotImplemented "
        f"for norb={norb}, nelec={nelec} (object: {obj!r})"
    )


def linear_operator(obj: Any, norb: int, nelec):
    """Return a LinearOperator for the object
    (python/ffsim/protocols/linear_operator_protocol.py:40-60)."""
    method = getattr(obj, "_linear_operator_", None)
    if method is not None:
        return method(norb=norb, nelec=nelec)
    raise TypeError(f"linear_operator: {type(obj)} defines no _linear_operator_.")


def approx_eq(obj: Any, other: Any, rtol: float = 1e-5, atol: float = 1e-8) -> bool:
    """Whether two objects are equal within ``rtol``/``atol`` (``np.isclose``
    sense).  Same dispatch as ffsim.approx_eq
    (python/ffsim/protocols/approximate_equality_protocol.py:40-70): ``obj._approx_eq_``,
    then ``other._approx_eq_``, each skipped when missing or NotImplemented,
    then ``==``."""
    for a, b in ((obj, other), (other, obj)):
        method = getattr(a, "_approx_eq_", None)
        if method is not None:
            result = method(b, rtol=rtol, atol=atol)
            if result is not NotImplemented:
                return result
    return obj == other


def diag(obj: Any, norb: int, nelec):
    """Diagonal entries of a linear operator in the determinant basis
    (python/ffsim/protocols/diagonal_protocol.py:40-50): ``obj._diag_``, else
    TypeError."""
    method = getattr(obj, "_diag_", None)
    if method is not None:
        return method(norb=norb, nelec=nelec)
    raise TypeError(
        f"diag: {type(obj)} defines no _diag_ that returns its diagonal."
    )


def trace(obj: Any, norb: int, nelec) -> float:
    """Trace of a linear operator (python/ffsim/protocols/trace_protocol.py:34-47):
    ``obj._trace_``, else the sum of ``obj._diag_``, else TypeError."""
    method = getattr(obj, "_trace_", None)
    if method is not None:
        return method(norb=norb, nelec=nelec)
    method = getattr(obj, "_diag_", None)
    if method is not None:
        return np.sum(method(norb=norb, nelec=nelec))
    raise TypeError(
        f"trace: {type(obj)} defines neither _trace_ nor a _diag_ that returns its "
        "diagonal."
    )
