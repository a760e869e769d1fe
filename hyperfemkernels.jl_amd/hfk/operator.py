"""Host mirror of the reference's operator surface, over the C ABI of libhfk.

Mirrors (src/assembly.jl:630-666, 904-933; src/matrixfree.jl:33-260):

* ``KernelOperator(form, U, V; jacobian=:assembled|:matrixfree, backend=:mi355x)`` -> ``KernelOperator(form,
  space, jacobian="assembled"|"matrixfree", backend="mi355x")``, with ``residual``/``jacobian``/
  ``residual_and_jacobian``/``allocate_residual``/``allocate_jacobian``/``zero_initial_guess``/
  ``sparsity_pattern``/``nonzero_positions``;
* ``MatrixFreeJacobian(op)`` with ``prepare``, ``mul`` (``mul!``), ``@`` (``*``), ``shape`` (``size``).

Julia's ``!`` methods overwrite their first argument and return it; the methods here do the same.  Errors map
to the reference's exceptions: ``ValueError`` for ``ArgumentError``/``DimensionMismatch``, ``RuntimeError``
for ``ErrorException``.  Everything runs in the HIP library: there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import LAWS, STORE, check, ptr


@dataclass(frozen=True)
class HyperelasticForm:
    """The weak form ``res(u, v) = ∫(∇(v)' ⊙ (P ∘ (F ∘ ∇(u)')))dΩ`` with ``jac(u, du, v) = ∫(∇(v)' ⊙ (dP ∘
    (F ∘ ∇(u)', ∇(du)')))dΩ`` (README.md:53-67) for a built-in law, ``F = I + ∇u'``, on a measure of
    ``degree`` (``Measure(Ω, degree)``: ``degree // 2 + 1`` Gauss points per direction)."""

    law: str = "svk"
    lam: float = 2.0
    mu: float = 1.0
    degree: int = 4

    @property
    def points1d(self) -> int:
        return self.degree // 2 + 1


class Context:
    """A HIP device and stream (``hfk_open``)."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        h = C.c_void_p()
        check(self.lib.hfk_open(int(device), C.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        check(self.lib.hfk_synchronize(self.handle), self.handle)

    def close(self):
        if self.handle:
            self.lib.hfk_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class KernelOperator:
    """Residual and Jacobian of ``form`` on ``space`` (a ``mesh.HexSpace``), assembled on the GPU."""

    def __init__(self, form: HyperelasticForm, space, jacobian: str = "assembled", backend: str = "mi355x",
                 dirichlet_values=None, ctx: Context | None = None, deterministic: bool = True):
        if backend != "mi355x":
            raise ValueError(f"HyperFEMKernels: unknown backend {backend!r} (this build is mi355x only)")
        if jacobian not in ("assembled", "matrixfree"):
            raise ValueError("HyperFEMKernels: jacobian must be 'assembled' or 'matrixfree'")
        if form.law not in LAWS:
            raise ValueError(f"HyperFEMKernels: unknown law {form.law!r}")
        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        self.form, self.space, self.jacobian_kind = form, space, jacobian
        self.nfree, self.ndir = int(space.nfree), int(space.ndir)
        h = C.c_void_p()
        cd = np.ascontiguousarray(space.cell_dofs, dtype=np.int32)
        ld = np.ascontiguousarray(space.ldof, dtype=np.int32)
        npm = np.ascontiguousarray(space.nodeperm, dtype=np.int32)
        xy = np.ascontiguousarray(space.coords, dtype=np.float64)
        check(self.lib.hfk_op_create(self.ctx.handle, space.Q, form.points1d, space.ncells, ptr(cd, C.c_int32),
                                     ptr(ld, C.c_int32), ptr(npm, C.c_int32), ptr(xy, C.c_double), self.nfree,
                                     self.ndir, LAWS[form.law], float(form.lam), float(form.mu), C.byref(h)),
              self.ctx.handle)
        self.handle = h
        self._pattern = None
        self.deterministic = True
        if not deterministic:
            self.set_deterministic(False)
        if dirichlet_values is not None:
            self.set_dirichlet(dirichlet_values)

    # --- data --------------------------------------------------------------------------------------------
    def set_refinement(self, coarse, parent=None, child=None):
        """Declare this operator's mesh a nested refinement of ``coarse``'s (an h-level pair of a p-multigrid):
        ``parent``/``child`` per cell (default ``mesh.refinement`` of the two spaces); ``coarse=None`` removes it."""
        if coarse is None:
            check(self.lib.hfk_op_set_refinement(self.handle, None, None, None), self.ctx.handle)
            return self
        from .mesh import refinement

        if parent is None:
            parent, child = refinement(self.space, coarse.space)
        pa = np.ascontiguousarray(parent, dtype=np.int32)
        ch = np.ascontiguousarray(child, dtype=np.int32)
        check(self.lib.hfk_op_set_refinement(self.handle, coarse.handle, ptr(pa, C.c_int32), ptr(ch, C.c_int32)),
              self.ctx.handle)
        self._refines = coarse  # kept alive while declared
        return self

    def set_dirichlet(self, dvals):
        d = np.ascontiguousarray(dvals, dtype=np.float64)
        if d.size != self.ndir:
            raise ValueError("HyperFEMKernels: Dirichlet values of the wrong length")
        check(self.lib.hfk_op_set_dirichlet(self.handle, ptr(d, C.c_double)), self.ctx.handle)
        self._dvals = d

    def update(self, form: HyperelasticForm):
        """``update(op, res, jac)`` (src/assembly.jl:693-744): the operator of ``form``.  With the structure of this
        operator's form (same law and quadrature) the numbers are rebound in place (``hfk_op_set_parameters``:
        colouring, pattern, tables and compiled laws kept) and ``self`` is returned; otherwise a new operator on the
        same space, with the same Jacobian kind and Dirichlet values, is built.  Prepared tangents are invalidated
        either way."""
        if form.law == self.form.law and form.degree == self.form.degree:
            check(self.lib.hfk_op_set_parameters(self.handle, float(form.lam), float(form.mu)), self.ctx.handle)
            self.form = form
            return self
        return KernelOperator(form, self.space, jacobian=self.jacobian_kind, dirichlet_values=getattr(self, "_dvals", None),
                              ctx=self.ctx)

    def set_traction(self, face_cell, face_local, traction):
        """A dead traction on faces of the cells (``mesh.boundary_faces``): the Neumann load -Λ ∫ t·v dΓ in the
        residual and -Λ ∫ t·u dΓ in the energy; ``traction`` (3,) for every face or (nfaces, 3); ``face_cell=None``
        removes it."""
        if face_cell is None:
            check(self.lib.hfk_op_set_traction(self.handle, 0, None, None, None), self.ctx.handle)
            return
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        fl = np.ascontiguousarray(face_local, dtype=np.int32)
        tr = np.ascontiguousarray(np.broadcast_to(np.asarray(traction, dtype=np.float64), (fc.size, 3)))
        if fl.size != fc.size:
            raise ValueError("HyperFEMKernels: DimensionMismatch: traction faces")
        check(self.lib.hfk_op_set_traction(self.handle, fc.size, ptr(fc, C.c_int32), ptr(fl, C.c_int32),
                                           ptr(tr, C.c_double)), self.ctx.handle)

    def set_robin(self, face_cell, face_local, stiffness):
        """A Robin (spring) term k ∫ u·v dΓ on faces of the cells: in residual!, Jv, the assembled Jacobian, the
        node blocks and the energy; ``stiffness`` a number or one per face; ``face_cell=None`` removes it.  Prepared
        tangents are invalidated (prepare again)."""
        if face_cell is None:
            check(self.lib.hfk_op_set_robin(self.handle, 0, None, None, None), self.ctx.handle)
            return
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        fl = np.ascontiguousarray(face_local, dtype=np.int32)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(stiffness, dtype=np.float64), (fc.size,)))
        if fl.size != fc.size:
            raise ValueError("HyperFEMKernels: DimensionMismatch: Robin faces")
        check(self.lib.hfk_op_set_robin(self.handle, fc.size, ptr(fc, C.c_int32), ptr(fl, C.c_int32),
                                        ptr(k, C.c_double)), self.ctx.handle)

    def set_skeleton(self, facets, source: str | None = None, name: str | None = None, params=(1.0,)):
        """Skeleton integrals on interior facets (``mesh.interior_facets``: plus cell, plus face, minus cell, minus
        face, orientation, each an array of the facets): the built-in penalty on the jump of the displacement
        gradient (gamma = params[0]) or a facet point form given as HIP source (``source``, struct ``name``).
        residual!, Jv, jacobian!, the energy and the sparsity pattern include them; ``facets=None`` removes them.
        Prepared tangents are invalidated (prepare again)."""
        self._pattern = None
        if facets is None:
            check(self.lib.hfk_op_set_skeleton(self.handle, 0, None, None, None, None, None, None, None, None, 0),
                  self.ctx.handle)
            self.nfacets = 0
            return
        pc, pf, mc, mf, ori = (np.ascontiguousarray(a, dtype=np.int32) for a in facets)
        if not (pc.size == pf.size == mc.size == mf.size == ori.size):
            raise ValueError("HyperFEMKernels: DimensionMismatch: skeleton facets")
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        check(self.lib.hfk_op_set_skeleton(self.handle, pc.size, ptr(pc, C.c_int32), ptr(pf, C.c_int32),
                                           ptr(mc, C.c_int32), ptr(mf, C.c_int32), ptr(ori, C.c_int32),
                                           source.encode() if source is not None else None,
                                           name.encode() if name is not None else None,
                                           ptr(p, C.c_double) if p.size else None, p.size), self.ctx.handle)
        self.nfacets = pc.size

    def set_boundary_form(self, face_cell, face_local, source: str | None = None, name: str | None = None,
                          params=(1.0,)):
        """A boundary group with a compiled point form on faces of the cells (``mesh.boundary_faces``): the built-in
        follower pressure (p = params[0]) or a form given as HIP source (``source``, struct ``name``); in residual!,
        Jv, jacobian! and the energy.  ``face_cell=None`` removes it.  Prepared tangents are invalidated."""
        if face_cell is None:
            check(self.lib.hfk_op_set_boundary_form(self.handle, 0, None, None, None, None, None, 0), self.ctx.handle)
            return
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        fl = np.ascontiguousarray(face_local, dtype=np.int32)
        if fl.size != fc.size:
            raise ValueError("HyperFEMKernels: DimensionMismatch: boundary faces")
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
        check(self.lib.hfk_op_set_boundary_form(self.handle, fc.size, ptr(fc, C.c_int32), ptr(fl, C.c_int32),
                                                source.encode() if source is not None else None,
                                                name.encode() if name is not None else None,
                                                ptr(p, C.c_double) if p.size else None, p.size), self.ctx.handle)

    def skeleton_colors(self) -> np.ndarray:
        """The colour (1-based) of every facet of the skeleton."""
        c = np.zeros(self.nfacets, dtype=np.int32)
        check(self.lib.hfk_op_skeleton_colors(self.handle, ptr(c, C.c_int32)), self.ctx.handle)
        return c

    def set_load_factor(self, load_factor: float):
        """The load factor Λ of the traction (a number of the form: nothing is recomputed)."""
        check(self.lib.hfk_op_set_load_factor(self.handle, float(load_factor)), self.ctx.handle)

    def traction_load(self):
        """The unscaled traction load vector: (f on the free ids, f on the Dirichlet ids)."""
        f = np.zeros(self.nfree + self.ndir)
        check(self.lib.hfk_op_traction_load(self.handle, ptr(f, C.c_double)), self.ctx.handle)
        return f[:self.nfree], f[self.nfree:]

    def set_mu_field(self, mu_points):
        if mu_points is None:
            check(self.lib.hfk_op_set_mu_field(self.handle, None), self.ctx.handle)
            return
        m = np.ascontiguousarray(mu_points, dtype=np.float64)
        if m.size != self.space.ncells * self.form.points1d ** 3:
            raise ValueError("HyperFEMKernels: mu field must have np values per cell")
        check(self.lib.hfk_op_set_mu_field(self.handle, ptr(m, C.c_double)), self.ctx.handle)

    def set_mu_function(self, source: str, name: str, params=None):
        """Shear modulus mu(x), a function of the physical coordinates given as HIP source
        ``__device__ double name(const double* x, const double* p)``, compiled with hiprtc and evaluated on the
        device at the quadrature points (the reference's function-of-x leaves, src/coordfunctions.jl:150-213);
        ``params``: the array ``p`` (may be None)."""
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
        check(self.lib.hfk_op_set_mu_function(self.handle, source.encode(), name.encode(),
                                              None if p is None or p.size == 0 else ptr(p, C.c_double),
                                              0 if p is None else int(p.size)), self.ctx.handle)

    def mu_field(self) -> np.ndarray:
        """The shear modulus at the quadrature points, (ncells, np) in tensor point order."""
        m = np.zeros((self.space.ncells, self.form.points1d ** 3))
        check(self.lib.hfk_op_mu_field(self.handle, ptr(m, C.c_double)), self.ctx.handle)
        return m

    def set_law(self, law):
        """Evaluate the residual, tangents and Jacobian with a law compiled at run time (``laws.CompiledLaw``;
        ``None``: the form's built-in law).  Prepared tangents are invalidated (prepare again)."""
        if law is not None and law.ctx is not self.ctx:
            raise ValueError("HyperFEMKernels: the law was compiled for another context")
        check(self.lib.hfk_op_set_law(self.handle, law.handle if law is not None else None), self.ctx.handle)
        self.law = law  # kept alive while the operator uses it
        return self

    def info(self) -> dict:
        a = np.zeros(8, dtype=np.int64)
        check(self.lib.hfk_op_info(self.handle, ptr(a, C.c_int64)), self.ctx.handle)
        keys = ["nfree", "ndir", "ncells", "nd", "np", "ncolors", "tangent_bytes", "nnz"]
        return dict(zip(keys, (int(v) for v in a)))

    def set_deterministic(self, deterministic: bool = True):
        """Products colour by colour (True: the reference's summation order, bitwise reproducible) or in one
        launch with FP64 atomic adds (False: faster; the order of the additions varies between runs)."""
        check(self.lib.hfk_op_set_deterministic(self.handle, int(bool(deterministic))), self.ctx.handle)
        self.deterministic = bool(deterministic)
        return self

    def colors(self) -> np.ndarray:
        c = np.zeros(self.space.ncells, dtype=np.int32)
        check(self.lib.hfk_op_colors(self.handle, ptr(c, C.c_int32)), self.ctx.handle)
        return c

    # --- Gridap Algebra interface ------------------------------------------------------------------------
    def _vec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.size != self.nfree:
            raise ValueError(f"HyperFEMKernels: DimensionMismatch: vector of length {x.size}, expected {self.nfree}")
        return x

    def allocate_residual(self, x=None):
        return np.zeros(self.nfree)

    def zero_initial_guess(self):
        return np.zeros(self.nfree)

    def residual(self, b, x):
        """``residual!(b, op, x)``: b overwritten with the residual at state x (free rows)."""
        x = self._vec(x)
        if b.dtype != np.float64 or not b.flags.c_contiguous or b.size != self.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: residual vector")
        check(self.lib.hfk_residual(self.handle, ptr(x, C.c_double), ptr(b, C.c_double)), self.ctx.handle)
        return b

    def pattern(self):
        if self._pattern is None:
            nnz = C.c_int64()
            check(self.lib.hfk_pattern_nnz(self.handle, C.byref(nnz)), self.ctx.handle)
            colptr = np.zeros(self.nfree + 1, dtype=np.int64)
            rowval = np.zeros(nnz.value, dtype=np.int64)
            nzpos = np.zeros((self.space.ncells, self.space.nd * self.space.nd), dtype=np.int32)
            check(self.lib.hfk_pattern(self.handle, ptr(colptr, C.c_int64), ptr(rowval, C.c_int64),
                                       ptr(nzpos, C.c_int32)), self.ctx.handle)
            self._pattern = (colptr, rowval, nzpos)
        return self._pattern

    def sparsity_pattern(self):
        """CSC pattern (1-based colptr/rowval as Julia's SparseMatrixCSC) as a scipy matrix of zeros."""
        import scipy.sparse as sp

        colptr, rowval, _ = self.pattern()
        return sp.csc_matrix((np.zeros(rowval.size), rowval - 1, colptr - 1), shape=(self.nfree, self.nfree))

    def nonzero_positions(self):
        return self.pattern()[2]

    def allocate_jacobian(self, x=None):
        return self.sparsity_pattern()

    def jacobian(self, A, x):
        """``jacobian!(A, op, x)``: the nonzeros of ``A`` (scipy CSC with this operator's pattern) overwritten."""
        x = self._vec(x)
        colptr, rowval, _ = self.pattern()
        if A.shape != (self.nfree, self.nfree) or A.nnz != rowval.size or not np.array_equal(A.indptr, colptr - 1):
            raise RuntimeError("HyperFEMKernels: jacobian! on a matrix with another sparsity pattern")
        if A.data.dtype != np.float64 or not A.data.flags.c_contiguous:
            A.data = np.ascontiguousarray(A.data, dtype=np.float64)
        check(self.lib.hfk_jacobian(self.handle, ptr(x, C.c_double), ptr(A.data, C.c_double)), self.ctx.handle)
        return A

    def jacobian_values(self, x):
        x = self._vec(x)
        colptr, rowval, _ = self.pattern()
        nz = np.zeros(rowval.size)
        check(self.lib.hfk_jacobian(self.handle, ptr(x, C.c_double), ptr(nz, C.c_double)), self.ctx.handle)
        return nz

    def residual_dev(self, b_ptr: int, x_ptr: int, stream: int = 0):
        """``residual!`` on device vectors (raw addresses of nfree doubles), enqueued on ``stream``."""
        check(self.lib.hfk_residual_dev(self.handle, C.c_void_p(x_ptr), C.c_void_p(b_ptr), C.c_void_p(stream)),
              self.ctx.handle)

    def jacobian_dev(self, nz_ptr: int, x_ptr: int, stream: int = 0):
        """``jacobian!`` into device CSC values (raw address of nnz doubles, the pattern of ``pattern()``)."""
        self.pattern()
        check(self.lib.hfk_jacobian_dev(self.handle, C.c_void_p(x_ptr), C.c_void_p(nz_ptr), C.c_void_p(stream)),
              self.ctx.handle)

    def residual_and_jacobian(self, b, A, x):
        self.residual(b, x)
        self.jacobian(A, x)
        return b, A

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hfk_op_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MatrixFreeJacobian:
    """``y = J(x) dx`` without assembling (src/matrixfree.jl:33-260).  ``storage`` selects how the point
    tangents are kept: "full64" (the reference's 81 FP64 per point), "sym64" (45, for symmetric tangents),
    "full32"/"sym32" (mixed precision)."""

    def __init__(self, op: KernelOperator, storage: str = "full64"):
        if storage not in STORE:
            raise ValueError(f"HyperFEMKernels: unknown tangent storage {storage!r}")
        self.op, self.storage = op, storage
        self.x = None
        self.shape = (op.nfree, op.nfree)

    @property
    def size(self):
        return self.shape

    def prepare(self, x):
        """``prepare!(mf, x)``: tangents at the state x (a copy of x is kept, as the reference does)."""
        x = self.op._vec(x)
        self.x = x.copy()
        check(self.op.lib.hfk_mf_prepare(self.op.handle, ptr(x, C.c_double), STORE[self.storage]),
              self.op.ctx.handle)
        return self

    def mul(self, y, dx):
        """``mul!(y, mf, dx)``."""
        dx = self.op._vec(dx)
        if y.dtype != np.float64 or not y.flags.c_contiguous or y.size != self.op.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: output vector")
        check(self.op.lib.hfk_mf_apply(self.op.handle, ptr(dx, C.c_double), ptr(y, C.c_double)), self.op.ctx.handle)
        return y

    def __matmul__(self, dx):
        return self.mul(np.zeros(self.op.nfree), dx)

    def tangents(self) -> np.ndarray:
        """Stored tangents in reference coordinates, (ncells, np, 81) with entry i + 9 j."""
        sp = self.op.space
        npts = self.op.form.points1d ** 3
        D = np.zeros((sp.ncells, npts, 81))
        check(self.op.lib.hfk_mf_tangents(self.op.handle, ptr(D, C.c_double)), self.op.ctx.handle)
        return D

    # device-pointer variants (torch tensors or raw addresses), asynchronous on `stream`
    def prepare_dev(self, x_ptr: int, stream: int = 0):
        check(self.op.lib.hfk_mf_prepare_dev(self.op.handle, C.c_void_p(x_ptr), STORE[self.storage],
                                             C.c_void_p(stream)), self.op.ctx.handle)

    def mul_dev(self, y_ptr: int, dx_ptr: int, stream: int = 0):
        check(self.op.lib.hfk_mf_apply_dev(self.op.handle, C.c_void_p(dx_ptr), C.c_void_p(y_ptr),
                                           C.c_void_p(stream)), self.op.ctx.handle)
