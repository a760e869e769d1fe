"""Verification of a device operator against a reference assembly (``check``, src/verify.jl:9-17).

The reference assembles with Gridap and with the kernels and returns the relative differences of the residual and
of the Jacobian.  Here the reference assembly is passed in by the caller, as functions of the free values (Gridap's
``Algebra.residual``/``Algebra.jacobian`` on a Julia host; any host assembly otherwise).
"""
from __future__ import annotations

from typing import Callable, NamedTuple

import numpy as np


class CheckResult(NamedTuple):
    res: float
    jac: float


def _norm(a) -> float:
    if hasattr(a, "toarray") or hasattr(a, "tocsc"):
        import scipy.sparse.linalg as spla

        return float(spla.norm(a))
    return float(np.linalg.norm(np.asarray(a)))


def check(op, residual_ref: Callable, jacobian_ref: Callable, x) -> CheckResult:
    """``check(res, jac, U, V, x; op)``: ``(res = |b_k - b_r| / |b_r|, jac = |A_k - A_r| / |A_r|)`` with b_k, A_k the
    residual and the assembled Jacobian of ``op`` at the free values ``x`` (a matrix-free operator assembles through
    its sparsity pattern, as the reference does) and b_r, A_r the reference's; Frobenius norms, absolute when the
    reference is zero."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    bk = op.residual(np.zeros(op.nfree), x)
    Ak = op.jacobian(op.allocate_jacobian(), x)
    br = np.asarray(residual_ref(x), dtype=np.float64)
    Ar = jacobian_ref(x)

    def rel(a, b):
        nb = _norm(b)
        d = _norm(a - b)
        return d if nb == 0 else d / nb

    return CheckResult(res=rel(bk, br), jac=rel(Ak, Ar))
