"""The Oseen-BDF2 step on tetrahedra (Taylor-Hood P2-P1, SUPG/PSPG/LSIC) as an operator with the reference's
Algebra interface: residual!, sparsity_pattern / nonzero_positions, allocate_jacobian, jacobian!
(src/assembly.jl:768-933), over libhfk's hfk_oseen_* entry points.  The weak form is that of test/common.jl:165-228:
volume terms, the elementwise Laplacian of un in the stabilisation load (``set_laplacian``) and the outlet terms
(``set_outlet``: backflow, Robin, Windkessel pressure).  The point data (advecting velocity w, stabilisation
velocity un, history f, Laplacian) are either given at the points by the host (``set_fields``, ``set_laplacian``)
or computed on the device from the FE functions of the two previous steps (``step_fields``).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, ptr
from .operator import Context, default_context


class OseenOperator:
    """``OseenOperator(space, rho, mu, dt; bdf2)``: space is a tetmesh.TetP2P1Space (or any object with the same
    fields); c0 = 1.5/dt (BDF2) or 1/dt (backward Euler)."""

    def __init__(self, space, rho: float = 1.0, mu: float = 0.01, dt: float = 0.01, bdf2: bool = True,
                 quadrature=None, dirichlet_values=None, ctx: Context | None = None):
        from .tetmesh import tet_quadrature_deg5

        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.space = space
        self.rho, self.mu, self.dt, self.bdf2 = float(rho), float(mu), float(dt), bool(bdf2)
        self.c0 = (1.5 if bdf2 else 1.0) / dt
        qp, qw = quadrature if quadrature is not None else tet_quadrature_deg5()
        self.qpoints = np.ascontiguousarray(qp, dtype=np.float64)
        self.qweights = np.ascontiguousarray(qw, dtype=np.float64)
        self.np = self.qweights.size
        self.nfree, self.ndir = int(space.nfree), int(space.ndir)
        cd = np.ascontiguousarray(space.cell_dofs, dtype=np.int32)
        xy = np.ascontiguousarray(space.coords, dtype=np.float64)
        # the space's local order, when it is not the library's own (tetmesh.reorder_local): the P2 and P1 table
        # node of each of the library's nodes and the local dof of each (component, table node)
        order = [getattr(space, k, None) for k in ("nodeperm", "pnodeperm", "ldof")]
        order = [None if a is None else np.ascontiguousarray(a, dtype=np.int32) for a in order]
        h = C.c_void_p()
        check(self.lib.hfk_oseen_create(self.ctx.handle, space.ncells, ptr(cd, C.c_int32),
                                        *[None if a is None else ptr(a, C.c_int32) for a in order],
                                        ptr(xy, C.c_double), self.np,
                                        ptr(self.qpoints, C.c_double), ptr(self.qweights, C.c_double), self.nfree,
                                        self.ndir, self.rho, self.mu, self.c0, C.byref(h)), self.ctx.handle)
        self.handle = h
        self._pattern = None
        self.form = None
        if dirichlet_values is not None:
            self.set_dirichlet(dirichlet_values)

    def set_form(self, form=None, params=None):
        """Evaluate the volume integrand with a point form compiled at run time (``formgen.OseenPointForm``, e.g.
        ``formgen.oseen_supg()``), its own parameters by name in ``params``; ``None`` restores the built-in
        integrand.  rho, mu and c0 stay the operator's (``update``)."""
        if form is None:
            check(self.lib.hfk_oseen_set_form(self.handle, None, None, None, 0), self.ctx.handle)
            self.form = None
            return self
        params = params or {}
        extra = np.ascontiguousarray([float(params[n]) for n in form.params], dtype=np.float64)
        src = form.source().encode()
        check(self.lib.hfk_oseen_set_form(self.handle, src, form.name.encode(),
                                          ptr(extra, C.c_double) if extra.size else None, int(extra.size)),
              self.ctx.handle)
        self.form = form
        return self

    def set_blocks(self, mask=None, cell_mask=None):
        """Field blocks that carry a term (src/assembly.jl:276-323, 825-893): ``mask`` 2 x 2 booleans [test field,
        trial field] (velocity 0, pressure 1) for every cell, or ``cell_mask`` (ncells, 2, 2); both None: every
        block.  The sparsity pattern is rebuilt at its next use."""
        m = None if mask is None else np.ascontiguousarray(np.asarray(mask, dtype=np.int32).T.reshape(-1))
        cm = None
        if cell_mask is not None:
            cm = np.ascontiguousarray(np.transpose(np.asarray(cell_mask, dtype=np.int32), (0, 2, 1)).reshape(-1))
        check(self.lib.hfk_oseen_set_blocks(self.handle, None if m is None else ptr(m, C.c_int32),
                                            None if cm is None else ptr(cm, C.c_int32)), self.ctx.handle)
        self._pattern = None
        return self

    def set_dirichlet(self, dvals):
        d = np.ascontiguousarray(dvals, dtype=np.float64)
        if d.size != self.ndir:
            raise ValueError("HyperFEMKernels: DimensionMismatch: Dirichlet values")
        check(self.lib.hfk_oseen_set_dirichlet(self.handle, ptr(d, C.c_double)), self.ctx.handle)

    def point_coordinates(self) -> np.ndarray:
        return self.space.quadrature_points(self.qpoints)

    def set_fields(self, w, un, f):
        """Point data (ncells, np, 3) each: advecting velocity, stabilisation velocity, history term."""
        arr = np.concatenate([np.asarray(a, dtype=np.float64).reshape(self.space.ncells, self.np, 3) for a in (w, un, f)],
                             axis=2)
        arr = np.ascontiguousarray(arr)
        check(self.lib.hfk_oseen_set_fields(self.handle, ptr(arr, C.c_double)), self.ctx.handle)

    def set_steady_advection(self, wfun):
        """The step of a steady advecting field w: un = un-1 = w, so w_adv = w, the stabilisation velocity is w
        and the BDF history is f = c0 w (c1 un - c2 un-1 with c1 - c2 = c0)."""
        wq = np.asarray(wfun(self.point_coordinates().reshape(-1, 3))).reshape(self.space.ncells, self.np, 3)
        self.set_fields(wq, wq, self.c0 * wq)

    def set_laplacian(self, lap=None):
        """Elementwise Laplacian of un at the points, (ncells, np, 3) or per cell (ncells, 3) (``tetmesh.
        p2_laplacian``); None removes the term from the stabilisation load."""
        if lap is None:
            check(self.lib.hfk_oseen_set_laplacian(self.handle, None), self.ctx.handle)
            return
        lap = np.asarray(lap, dtype=np.float64)
        if lap.shape == (self.space.ncells, 3):
            lap = np.repeat(lap[:, None, :], self.np, axis=1)
        if lap.shape != (self.space.ncells, self.np, 3):
            raise ValueError("HyperFEMKernels: DimensionMismatch: Laplacian point data")
        lap = np.ascontiguousarray(lap)
        check(self.lib.hfk_oseen_set_laplacian(self.handle, ptr(lap, C.c_double)), self.ctx.handle)

    def step_coefficients(self):
        """(aw, bw, c1, c2) of the step (test/common.jl:185-190): w = aw un + bw unm1 and f = c1 un - c2 unm1; BDF2
        (2, -1, 2/dt, 1/(2 dt)) or backward Euler (1, 0, 1/dt, 0)."""
        if self.bdf2:
            return 2.0, -1.0, 2.0 / self.dt, 0.5 / self.dt
        return 1.0, 0.0, 1.0 / self.dt, 0.0

    def step_fields(self, un, un_dir, unm1, unm1_dir, laplacian: bool = True, coefficients=None):
        """The point data of the step computed on the device from the FE functions un = u^n and unm1 = u^{n-1}:
        free values (nfree; pressure entries not read) and Dirichlet values (ndir) of each.  Sets w, un, f at the
        points, the Laplacian of un (``laplacian``; False removes the term) and w at the outlet backflow points;
        ``coefficients`` (aw, bw, c1, c2) default to ``step_coefficients()``."""
        aw, bw, c1, c2 = coefficients if coefficients is not None else self.step_coefficients()
        a = [np.ascontiguousarray(v, dtype=np.float64) for v in (un, un_dir, unm1, unm1_dir)]
        if a[0].size != self.nfree or a[2].size != self.nfree or a[1].size != self.ndir or a[3].size != self.ndir:
            raise ValueError("HyperFEMKernels: DimensionMismatch: FE-function values")
        check(self.lib.hfk_oseen_step_fields(self.handle, *[ptr(v, C.c_double) for v in a], float(aw), float(bw),
                                             float(c1), float(c2), int(bool(laplacian))), self.ctx.handle)

    def step_fields_dev(self, un_ptr: int, un_dir_ptr: int, unm1_ptr: int, unm1_dir_ptr: int, laplacian: bool = True,
                        coefficients=None, stream: int = 0):
        """``step_fields`` with device pointers (torch tensors' data_ptr() or raw addresses), asynchronous on
        ``stream``."""
        aw, bw, c1, c2 = coefficients if coefficients is not None else self.step_coefficients()
        check(self.lib.hfk_oseen_step_fields_dev(self.handle, C.c_void_p(un_ptr), C.c_void_p(un_dir_ptr),
                                                 C.c_void_p(unm1_ptr), C.c_void_p(unm1_dir_ptr), float(aw), float(bw),
                                                 float(c1), float(c2), int(bool(laplacian)), C.c_void_p(stream)),
              self.ctx.handle)

    def set_outlet(self, faces, w=None, Pc: float = 0.0, ZA: float = 0.0):
        """Outlet terms on ``faces`` (``tetmesh.OutletFaces``): backflow -rho negpart(w.n) u.v, Robin ZA (u.n)(v.n),
        pressure load -Pc v.n (test/common.jl:212-213, 219).  ``w``: the advecting velocity at the backflow points,
        (nfaces, nqb, 3) or a callable of the physical coordinates; None: no backflow.  ``faces=None`` removes the
        terms."""
        if faces is None:
            check(self.lib.hfk_oseen_set_outlet(self.handle, 0, None, None, 1, None, 1, None, 0.0, 0.0),
                  self.ctx.handle)
            return
        nf, nqb = faces.xb.shape[0], faces.xb.shape[1]
        if w is None:
            wb = np.zeros((nf, nqb, 3))
        elif callable(w):
            wb = np.asarray(w(faces.xb_phys.reshape(-1, 3)), dtype=np.float64).reshape(nf, nqb, 3)
        else:
            wb = np.asarray(w, dtype=np.float64).reshape(nf, nqb, 3)
        xb = np.ascontiguousarray(np.concatenate([faces.xb, wb], axis=2))
        xo = np.ascontiguousarray(faces.xo, dtype=np.float64)
        cell = np.ascontiguousarray(faces.cell, dtype=np.int32)
        nrm = np.ascontiguousarray(faces.normal, dtype=np.float64)
        check(self.lib.hfk_oseen_set_outlet(self.handle, nf, ptr(cell, C.c_int32), ptr(nrm, C.c_double),
                                            xo.shape[1], ptr(xo, C.c_double), nqb, ptr(xb, C.c_double), float(Pc),
                                            float(ZA)), self.ctx.handle)
        self.outlet = (faces, xb)

    def set_deterministic(self, deterministic: bool = True):
        """Colour by colour with plain adds (True, bitwise reproducible) or one launch with FP64 atomic adds."""
        check(self.lib.hfk_oseen_set_deterministic(self.handle, int(bool(deterministic))), self.ctx.handle)
        return self

    def update(self, rho: float | None = None, mu: float | None = None, dt: float | None = None,
               bdf2: bool | None = None):
        """``update(op, res, jac)`` of the step with other numbers (test/test_cfd.jl:47-75): density, viscosity,
        time step, or the switch from backward Euler to BDF2 (c0 = 1/dt -> 1.5/dt, and the step coefficients of
        ``step_fields``); colouring, pattern, point data and outlet data are kept.  Returns self."""
        self.rho = self.rho if rho is None else float(rho)
        self.mu = self.mu if mu is None else float(mu)
        self.dt = self.dt if dt is None else float(dt)
        self.bdf2 = self.bdf2 if bdf2 is None else bool(bdf2)
        self.c0 = (1.5 if self.bdf2 else 1.0) / self.dt
        check(self.lib.hfk_oseen_set_parameters(self.handle, self.rho, self.mu, self.c0), self.ctx.handle)
        return self

    def set_outlet_pressure(self, Pc: float):
        check(self.lib.hfk_oseen_set_outlet_pressure(self.handle, float(Pc)), self.ctx.handle)

    def info(self) -> dict:
        a = np.zeros(6, dtype=np.int64)
        check(self.lib.hfk_oseen_info(self.handle, ptr(a, C.c_int64)), self.ctx.handle)
        return dict(zip(["nfree", "ndir", "ncells", "np", "ncolors", "nnz"], (int(v) for v in a)))

    def colors(self) -> np.ndarray:
        c = np.zeros(self.space.ncells, dtype=np.int32)
        check(self.lib.hfk_oseen_colors(self.handle, ptr(c, C.c_int32)), self.ctx.handle)
        return c

    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: state vector")
        return x

    def residual(self, b, x):
        x = self._vec(x)
        if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != self.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: residual vector")
        check(self.lib.hfk_oseen_residual(self.handle, ptr(x, C.c_double), ptr(b, C.c_double)), self.ctx.handle)
        return b

    def pattern_nnz(self) -> int:
        """Number of stored entries of the CSC pattern (built on the device at the first call)."""
        nnz = C.c_int64()
        check(self.lib.hfk_oseen_pattern_nnz(self.handle, C.byref(nnz)), self.ctx.handle)
        return int(nnz.value)

    def pattern(self):
        if self._pattern is None:
            nnz = C.c_int64()
            check(self.lib.hfk_oseen_pattern_nnz(self.handle, C.byref(nnz)), self.ctx.handle)
            colptr = np.zeros(self.nfree + 1, dtype=np.int64)
            rowval = np.zeros(nnz.value, dtype=np.int64)
            nzpos = np.zeros((self.space.ncells, 34 * 34), dtype=np.int32)
            check(self.lib.hfk_oseen_pattern(self.handle, ptr(colptr, C.c_int64), ptr(rowval, C.c_int64),
                                             ptr(nzpos, C.c_int32)), self.ctx.handle)
            self._pattern = (colptr, rowval, nzpos)
        return self._pattern

    def allocate_jacobian(self, x=None):
        import scipy.sparse as sp

        colptr, rowval, _ = self.pattern()
        return sp.csc_matrix((np.zeros(rowval.size), rowval - 1, colptr - 1), shape=(self.nfree, self.nfree))

    def jacobian(self, A, x=None):
        colptr, rowval, _ = self.pattern()
        if A.shape != (self.nfree, self.nfree) or A.nnz != rowval.size or not np.array_equal(A.indptr, colptr - 1):
            raise RuntimeError("HyperFEMKernels: jacobian! on a matrix with another sparsity pattern")
        if A.data.dtype != np.float64 or not A.data.flags.c_contiguous:
            A.data = np.ascontiguousarray(A.data, dtype=np.float64)
        check(self.lib.hfk_oseen_jacobian(self.handle, None, ptr(A.data, C.c_double)), self.ctx.handle)
        return A

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hfk_oseen_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
