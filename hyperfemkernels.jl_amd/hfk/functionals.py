"""Scalar functionals of the unknown evaluated by libhfk's kernels (KernelFunctional, src/functionals.jl:9-154):
the strain energy of the operator's law (reduce = "sum", the energy of a Newton line search) and the extremum over
the quadrature points of a pointwise function of the displacement gradient and the gradient of an increment
(reduce = "min" / "max", the injectivity step of a line search).  The increment is data of the functional, like the
reference's FE function ``duh`` whose values change in place between calls (test/solvers/test_functionals.jl:59-78).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, ptr


class KernelFunctional:
    """``KernelFunctional(op, reduce="sum")``: the strain energy sum_cells sum_q W(F) dV of ``op``'s law (built-in, or
    a compiled law that defines ``energy()``), ``fun(x) -> float``.

    ``KernelFunctional(op, reduce="min" | "max", source=..., name=..., params=None)``: the extremum over the
    quadrature points of ``g(Gu, Gdu, p)`` (HIP source ``__device__ double name(const double* Gu, const double* Gdu,
    const double* p)``, G[c + 3J] = du_c/dx_J), with the increment set by ``set_increment(dx, dx_dir)`` (its free
    and Dirichlet values; may change between calls)."""

    def __init__(self, op, reduce: str = "sum", source: str | None = None, name: str | None = None, params=None):
        if reduce not in ("sum", "min", "max"):
            raise ValueError("HyperFEMKernels: KernelFunctional: reduce must be sum, min or max")
        self.op, self.reduce = op, reduce
        self.lib, self.ctx = op.lib, op.ctx
        self.fn = None
        self.params = None if params is None else np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        self.dx = np.zeros(op.nfree)
        self.dx_dir = None
        if reduce == "sum":
            if source is not None:
                raise ValueError("HyperFEMKernels: KernelFunctional: the energy is the operator's law's")
            return
        if source is None or name is None:
            raise ValueError("HyperFEMKernels: KernelFunctional: reduce = min/max needs a pointwise function")
        h = C.c_void_p()
        check(self.lib.hfk_pointfn_create(self.ctx.handle, source.encode(), name.encode(), C.byref(h)), self.ctx.handle)
        self.fn = h

    def set_increment(self, dx, dx_dir=None):
        """The increment du: free values dx and Dirichlet values dx_dir (None: zero)."""
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        if dx.size != self.op.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: increment")
        self.dx = dx
        if dx_dir is not None:
            dx_dir = np.ascontiguousarray(dx_dir, dtype=np.float64)
            if dx_dir.size != self.op.ndir:
                raise ValueError("HyperFEMKernels: DimensionMismatch: increment Dirichlet values")
        self.dx_dir = dx_dir
        return self

    def __call__(self, x) -> float:
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.op.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: state vector")
        out = np.zeros(1)
        if self.reduce == "sum":
            check(self.lib.hfk_energy(self.op.handle, ptr(x, C.c_double), ptr(out, C.c_double)), self.ctx.handle)
            return float(out[0])
        p = self.params
        check(self.lib.hfk_extremum(self.op.handle, self.fn, ptr(x, C.c_double), ptr(self.dx, C.c_double),
                                    None if self.dx_dir is None else ptr(self.dx_dir, C.c_double),
                                    None if p is None or p.size == 0 else ptr(p, C.c_double),
                                    0 if p is None else int(p.size), 0 if self.reduce == "min" else 1,
                                    ptr(out, C.c_double)), self.ctx.handle)
        return float(out[0])

    def energy_dev(self, x_ptr: int, out_ptr: int, stream: int = 0):
        """The energy with device pointers (x and a one-double result), asynchronous on ``stream``."""
        check(self.lib.hfk_energy_dev(self.op.handle, C.c_void_p(x_ptr), C.c_void_p(out_ptr), C.c_void_p(stream)),
              self.ctx.handle)

    def __del__(self):
        try:
            if self.fn:
                self.lib.hfk_pointfn_destroy(self.fn)
        except Exception:
            pass
