"""Host mirror of the reference's solver layer over libhfk (all work on the GPU).

Mirrors src/preconditioners.jl (``BlockJacobi``, ``eigenvalue_bounds``), src/smoothers.jl (``Chebyshev``,
``smooth!``), src/transfers.jl (``Prolongation``, ``prolongate!``, ``restrict!``, ``inject!``),
src/pmultigrid.jl (``PMultigrid``) and src/solvers.jl (``MatrixFreeSolver``).  Numerical failure is reported
through status fields, never raised (src/solvers.jl:114-121); misuse raises.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import STORE, check, ptr
from .operator import KernelOperator, MatrixFreeJacobian


def _vec(a, n):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != n:
        raise ValueError(f"HyperFEMKernels: DimensionMismatch: vector of length {a.size}, expected {n}")
    return a


class BlockJacobi:
    """Inverse of the nodal 3x3 block diagonal of the prepared tangent (src/preconditioners.jl:100-342)."""

    def __init__(self, mf: MatrixFreeJacobian):
        self.op = mf.op
        self.ok = False
        self.update(mf)

    def update(self, mf: MatrixFreeJacobian):
        ok = C.c_int32()
        check(self.op.lib.hfk_block_jacobi_update(self.op.handle, C.byref(ok)), self.op.ctx.handle)
        self.ok = bool(ok.value)
        return self

    def ldiv(self, y, r):
        r = _vec(r, self.op.nfree)
        check(self.op.lib.hfk_block_jacobi_apply(self.op.handle, ptr(r, C.c_double), ptr(y, C.c_double)),
              self.op.ctx.handle)
        return y

    def blocks(self):
        nb = C.c_int64()
        lib = self.op.lib
        check(lib.hfk_block_jacobi_blocks(self.op.handle, C.byref(nb), None, None), self.op.ctx.handle)
        dofs = np.zeros((nb.value, 3), dtype=np.int32)
        Binv = np.zeros((nb.value, 9))
        check(lib.hfk_block_jacobi_blocks(self.op.handle, C.byref(nb), ptr(dofs, C.c_int32), ptr(Binv, C.c_double)),
              self.op.ctx.handle)
        return dofs, Binv


def eigenvalue_bounds(mf: MatrixFreeJacobian, k: int = 10):
    """Extreme real parts of the Ritz values of J M^-1 after k Arnoldi steps from chebyshev_start (the nodal
    blocks of the operator must be up to date)."""
    lo, hi = C.c_double(), C.c_double()
    check(mf.op.lib.hfk_eigenvalue_bounds(mf.op.handle, int(k), C.byref(lo), C.byref(hi)), mf.op.ctx.handle)
    return lo.value, hi.value


@dataclass(frozen=True)
class Chebyshev:
    """Smoother configuration (src/smoothers.jl:21-33)."""

    degree: int = 2
    range: float = 20.0
    inflate: float = 1.1
    eig_iters: int = 10

    def __post_init__(self):
        if self.degree < 1:
            raise ValueError("HyperFEMKernels: Chebyshev: degree must be >= 1")
        if not self.range > 1:
            raise ValueError("HyperFEMKernels: Chebyshev: range must be > 1")
        if not self.inflate >= 1:
            raise ValueError("HyperFEMKernels: Chebyshev: inflate must be >= 1")
        if self.eig_iters < 2:
            raise ValueError("HyperFEMKernels: Chebyshev: eig_iters must be >= 2")

    def damping(self) -> float:
        """max |p_k| on the interval: 1/T_k(theta/delta) (chebyshev_damping, src/smoothers.jl:36-40)."""
        b, a = 1.0, 1.0 / self.range
        th, de = (b + a) / 2, (b - a) / 2
        return 1.0 / np.cosh(self.degree * np.arccosh(th / de))


class ChebyshevState:
    """smoother_setup / smoother_update! / smooth! on one operator."""

    def __init__(self, cfg: Chebyshev, mf: MatrixFreeJacobian):
        self.cfg, self.op = cfg, mf.op
        h = C.c_void_p()
        check(self.op.lib.hfk_cheb_create(self.op.handle, cfg.degree, cfg.range, cfg.inflate, cfg.eig_iters,
                                          C.byref(h)), self.op.ctx.handle)
        self.handle = h
        self.update(mf)

    def update(self, mf: MatrixFreeJacobian):
        st = np.zeros(2, dtype=np.int32)
        iv = np.zeros(4)
        check(self.op.lib.hfk_cheb_update(self.handle, ptr(st, C.c_int32), ptr(iv, C.c_double)), self.op.ctx.handle)
        self.ok, self.indefinite = bool(st[0]), bool(st[1])
        self.lmin, self.lmax, self.a, self.b = (float(v) for v in iv)
        return self

    def smooth(self, x, rhs, zero_guess=False, residual=None):
        rhs = _vec(rhs, self.op.nfree)
        if x.dtype != np.float64 or not x.flags.c_contiguous or x.size != self.op.nfree:
            raise ValueError("HyperFEMKernels: DimensionMismatch: smoother iterate")
        rp = None if residual is None else ptr(residual, C.c_double)
        check(self.op.lib.hfk_cheb_smooth(self.handle, ptr(x, C.c_double), ptr(rhs, C.c_double), int(bool(zero_guess)),
                                          rp), self.op.ctx.handle)
        return x

    def __del__(self):
        try:
            self.op.lib.hfk_cheb_destroy(self.handle)
        except Exception:
            pass


class Prolongation:
    """Exact prolongation from the coarse operator's space to the fine one's on the same mesh."""

    def __init__(self, coarse: KernelOperator, fine: KernelOperator):
        h = C.c_void_p()
        check(fine.lib.hfk_transfer_create(coarse.handle, fine.handle, C.byref(h)), fine.ctx.handle)
        self.handle, self.coarse, self.fine, self.lib = h, coarse, fine, fine.lib
        self.shape = (fine.nfree, coarse.nfree)

    def prolongate(self, yf, xc):
        xc = _vec(xc, self.coarse.nfree)
        check(self.lib.hfk_prolongate(self.handle, ptr(xc, C.c_double), ptr(yf, C.c_double)), self.fine.ctx.handle)
        return yf

    def restrict(self, yc, yf):
        yf = _vec(yf, self.fine.nfree)
        check(self.lib.hfk_restrict(self.handle, ptr(yf, C.c_double), ptr(yc, C.c_double)), self.fine.ctx.handle)
        return yc

    def inject(self, xf, df):
        xf = _vec(xf, self.fine.nfree)
        df = _vec(df, self.fine.ndir) if self.fine.ndir else np.zeros(1)
        xc = np.zeros(self.coarse.nfree)
        dc = np.zeros(max(self.coarse.ndir, 1))
        check(self.lib.hfk_inject(self.handle, ptr(xf, C.c_double), ptr(df, C.c_double), ptr(xc, C.c_double),
                                  ptr(dc, C.c_double)), self.fine.ctx.handle)
        return xc, dc[: self.coarse.ndir]

    def __matmul__(self, xc):
        return self.prolongate(np.zeros(self.fine.nfree), xc)

    def __del__(self):
        try:
            self.lib.hfk_transfer_destroy(self.handle)
        except Exception:
            pass


class PMultigrid:
    """One V-cycle of p-multigrid as a preconditioner (src/pmultigrid.jl:223-343).  ``levels``: the operators,
    finest first (the first is the Newton's, matrix-free; the last assembled)."""

    def __init__(self, levels, smoother: Chebyshev = Chebyshev(), coarse_rtol: float = 1e-2,
                 coarse_maxiter: int = 500, smoother_storage: str | None = None, coarse_every: int = 1,
                 coarse_precision: str = "fp64", coarse: str = "pcg"):
        if len(levels) < 2:
            raise ValueError("HyperFEMKernels: PMultigrid needs at least two levels")
        if coarse_every < 1:
            raise ValueError("HyperFEMKernels: PMultigrid: coarse_every must be >= 1")
        self.levels = list(levels)
        arr = (C.c_void_p * len(levels))(*[op.handle.value for op in levels])
        h = C.c_void_p()
        lib = levels[0].lib
        sst = -1 if smoother_storage is None else STORE[smoother_storage]
        # coarse: "pcg" (block-Jacobi PCG to coarse_rtol) or "direct" (Cholesky of a small coarsest level, the
        # bottom of an h-hierarchy: a fixed cycle)
        if coarse not in ("pcg", "direct"):
            raise ValueError("HyperFEMKernels: PMultigrid: coarse must be 'pcg' or 'direct'")
        check(lib.hfk_pmg_create(arr, len(levels), smoother.degree, smoother.range, smoother.inflate,
                                 smoother.eig_iters, sst, 0 if coarse == "pcg" else 1, float(coarse_rtol),
                                 int(coarse_maxiter), C.byref(h)),
              levels[0].ctx.handle)
        self.handle, self.lib, self.smoother = h, lib, smoother
        self.ok = self.indefinite = False
        self.coarse_every = int(coarse_every)
        if coarse_every != 1:
            check(lib.hfk_pmg_set_coarse_every(h, int(coarse_every)), levels[0].ctx.handle)
        if coarse_precision not in ("fp64", "fp32"):
            raise ValueError("HyperFEMKernels: PMultigrid: coarse_precision must be fp64 or fp32")
        if coarse_precision == "fp32":  # the coarse PCG's products on FP32 values, FP64 vectors and accumulation
            check(lib.hfk_pmg_set_coarse_precision(h, 32), levels[0].ctx.handle)

    def coarse_builds(self) -> int:
        """Coarse set-ups so far (the reference's ``nbuild``)."""
        n = C.c_int64()
        check(self.lib.hfk_pmg_coarse_builds(self.handle, C.byref(n)), self.levels[0].ctx.handle)
        return int(n.value)

    def set_fallback(self, window: int = 16, factor: float = 0.5):
        """With a coarse network attached: set the network aside for the rest of a solve whose residual has not
        fallen by ``factor`` over the last ``window`` iterations (the plain cycle is the fallback); window 0: never."""
        check(self.lib.hfk_pmg_set_fallback(self.handle, int(window), float(factor)), self.levels[0].ctx.handle)
        return self

    def fallbacks(self) -> int:
        """Solves so far that fell back to the plain cycle."""
        n = C.c_int64()
        check(self.lib.hfk_pmg_fallbacks(self.handle, C.byref(n)), self.levels[0].ctx.handle)
        return int(n.value)

    def update(self):
        """preconditioner_update!: follow the state of the working operator's last prepare."""
        st = np.zeros(2, dtype=np.int32)
        check(self.lib.hfk_pmg_update(self.handle, ptr(st, C.c_int32)), self.levels[0].ctx.handle)
        self.ok, self.indefinite = bool(st[0]), bool(st[1])
        return self

    def ldiv(self, z, r):
        r = _vec(r, self.levels[0].nfree)
        check(self.lib.hfk_pmg_vcycle(self.handle, ptr(r, C.c_double), ptr(z, C.c_double)), self.levels[0].ctx.handle)
        return z

    def info(self) -> dict:
        a = np.zeros(6 + 6 * len(self.levels))
        check(self.lib.hfk_pmg_info(self.handle, ptr(a, C.c_double)), self.levels[0].ctx.handle)
        lv = [dict(lambda_min=a[6 + 6 * l], lambda_max=a[7 + 6 * l], a=a[8 + 6 * l], b=a[9 + 6 * l],
                   ok=bool(a[10 + 6 * l]), blocks_ok=bool(a[11 + 6 * l])) for l in range(len(self.levels) - 1)]
        return dict(ok=bool(a[1]), cycles=int(a[2]), last_coarse_iters=int(a[3]), coarse_iters=int(a[4]),
                    coarse_ok=bool(a[5]), levels=lv)

    def coarse_matvec(self, x, scalar_csc: bool = False):
        """A_c x with the assembled coarse matrix of the last update (the solver's block product, or the scalar CSC
        product with ``scalar_csc``)."""
        n = self.levels[-1].nfree
        x = _vec(x, n)
        y = np.zeros(n)
        check(self.lib.hfk_pmg_coarse_matvec(self.handle, ptr(x, C.c_double), ptr(y, C.c_double), int(bool(scalar_csc))),
              self.levels[0].ctx.handle)
        return y

    def __del__(self):
        try:
            self.lib.hfk_pmg_destroy(self.handle)
        except Exception:
            pass


@dataclass
class SolveResult:
    converged: bool
    iterations: int
    indefinite: bool
    residual: float
    rhs_norm: float
    fell_back: bool = False  # the coarse network was set aside for the plain cycle during the solve


KRYLOV = {"cg": 0, "fcg": 1, "gmres": 2, "fgmres": 3}


class MatrixFreeSolver:
    """Krylov solve on a prepared MatrixFreeJacobian (src/solvers.jl:33-174): ``krylov`` "gmres" (the reference's
    default), "fgmres", "cg" or "fcg" (flexible CG); GMRES and FGMRES restart every ``restart`` iterations;
    ``preconditioner`` None, "blockjacobi" or a PMultigrid."""

    def __init__(self, krylov: str = "gmres", rtol: float = 1e-8, atol: float = 0.0, maxiter: int = 500,
                 preconditioner=None, restart: int = 50):
        if krylov not in KRYLOV:
            raise ValueError(f"HyperFEMKernels: MatrixFreeSolver: krylov must be one of {sorted(KRYLOV)} (got {krylov!r})")
        if not (0 <= rtol < 1) or atol < 0:
            raise ValueError("HyperFEMKernels: MatrixFreeSolver: rtol must be in [0, 1) and atol >= 0")
        if maxiter <= 0 or restart <= 0:
            raise ValueError("HyperFEMKernels: MatrixFreeSolver: maxiter and restart must be positive")
        self.krylov, self.rtol, self.atol, self.maxiter, self.preconditioner = krylov, rtol, atol, maxiter, preconditioner
        self.restart = restart
        self.last = None

    def solve(self, x, mf: MatrixFreeJacobian, b) -> SolveResult:
        if not isinstance(mf, MatrixFreeJacobian):  # symbolic_setup on a matrix, src/solvers.jl:66-67
            raise RuntimeError("HyperFEMKernels: MatrixFreeSolver applies a MatrixFreeJacobian; build the "
                               "KernelOperator with jacobian = 'matrixfree'")
        op = mf.op
        b = _vec(b, op.nfree)
        pc = self.preconditioner
        prec, g = (0, None) if pc is None else ((1, None) if pc == "blockjacobi" else (2, pc.handle))
        st = np.zeros(3, dtype=np.int32)
        rn = np.zeros(2)
        n0 = pc.fallbacks() if prec == 2 else 0
        check(op.lib.hfk_krylov_solve(op.handle, g, KRYLOV[self.krylov], prec, ptr(b, C.c_double), ptr(x, C.c_double),
                                      self.rtol, self.atol, self.maxiter, self.restart, ptr(st, C.c_int32),
                                      ptr(rn, C.c_double)), op.ctx.handle)
        self.last = SolveResult(bool(st[0]), int(st[1]), bool(st[2]), float(rn[0]), float(rn[1]),
                                prec == 2 and pc.fallbacks() > n0)
        return self.last

    def solve_dev(self, x_ptr: int, mf: MatrixFreeJacobian, b_ptr: int, stream: int = 0) -> SolveResult:
        """The same on device vectors (raw addresses of nfree doubles); synchronises the stream at the end."""
        op = mf.op
        pc = self.preconditioner
        prec, g = (0, None) if pc is None else ((1, None) if pc == "blockjacobi" else (2, pc.handle))
        st = np.zeros(3, dtype=np.int32)
        rn = np.zeros(2)
        n0 = pc.fallbacks() if prec == 2 else 0
        check(op.lib.hfk_krylov_solve_dev(op.handle, g, KRYLOV[self.krylov], prec, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                          self.rtol, self.atol, self.maxiter, self.restart, ptr(st, C.c_int32),
                                          ptr(rn, C.c_double), C.c_void_p(stream)), op.ctx.handle)
        self.last = SolveResult(bool(st[0]), int(st[1]), bool(st[2]), float(rn[0]), float(rn[1]),
                                prec == 2 and pc.fallbacks() > n0)
        return self.last
