"""TEST INFRASTRUCTURE ONLY — ctypes binding of the oracle (liboracle.so, built by oracle/Makefile).

The oracle is the CPU restatement of the reference's FP64 algorithms (hfk_oracle.cpp, each function citing the
HyperFEMKernels.jl file:line it follows).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this module, and only as the checker: the product path (hfk, libhfk.so) never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
_lib = None

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    L = C.CDLL(LIB)
    sig = {
        "orc_gauss01": (None, [C.c_int, _f64p, _f64p]),
        "orc_lagrange1d": (None, [C.c_int, C.c_double, _f64p, _f64p]),
        "orc_law_P": (None, [C.c_int, C.c_double, C.c_double, _f64p, _f64p]),
        "orc_law_T": (None, [C.c_int, C.c_double, C.c_double, _f64p, _f64p]),
        "orc_law_W": (C.c_double, [C.c_int, C.c_double, C.c_double, _f64p]),
        "orc_op_create": (C.c_void_p, [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p, C.c_int64, C.c_int64,
                                       C.c_int, C.c_double, C.c_double, _f64p]),
        "orc_op_destroy": (None, [C.c_void_p]),
        "orc_op_set_mu_field": (None, [C.c_void_p, _f64p]),
        "orc_op_set_dirichlet": (None, [C.c_void_p, _f64p]),
        "orc_op_ncolors": (C.c_int, [C.c_void_p]),
        "orc_op_colors": (C.c_int, [C.c_void_p, _i32p]),
        "orc_last_error": (C.c_char_p, []),
        "orc_prepare": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p, _f64p]),
        "orc_jv": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p, _f64p, _f64p, C.c_int]),
        "orc_energy": (C.c_double, [C.c_void_p, _f64p]),
        "orc_point_gradients": (C.c_int, [C.c_void_p, _f64p, _f64p]),
        "orc_face_load": (C.c_int, [C.c_void_p, C.c_int64, _i32p, _i32p, _f64p, _f64p]),
        "orc_face_mass": (C.c_int, [C.c_void_p, C.c_int64, _i32p, _i32p, _f64p, _f64p]),
        "orc_residual": (C.c_int, [C.c_void_p, _f64p, _f64p]),
        "orc_element_jacobians": (C.c_int, [C.c_void_p, _f64p, _f64p]),
        "orc_pattern_nnz": (C.c_int64, [C.c_void_p]),
        "orc_pattern": (C.c_int, [C.c_void_p, _i64p, _i64p, _i32p]),
        "orc_jacobian": (C.c_int, [C.c_void_p, _f64p, _i32p, _f64p, C.c_int64]),
        "orc_node_blocks": (C.c_int, [C.c_void_p, _i32p, _i32p, C.c_int64, _i64p]),
        "orc_blocks_from_tangents": (C.c_int, [C.c_void_p, _f64p, _f64p, _f64p, _i32p, _i32p, C.c_int64, _f64p]),
        "orc_apply_blocks": (None, [_f64p, _i32p, C.c_int64, _f64p, _f64p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


LAWS = {"svk": 0, "neohookean": 1, "linear": 2}


def gauss01(P):
    x, w = np.zeros(P), np.zeros(P)
    load().orc_gauss01(P, _p(x, C.c_double), _p(w, C.c_double))
    return x, w


def physical_points(coords, P):
    """Physical coordinates (ncells, P^3, 3) of the tensor Gauss points q = a + P b + P^2 c under the trilinear maps
    of the cells' vertices coords (ncells, 8, 3), vertex v = i + 2j + 4k (x = sum_v M_v(xi) x_v, src/assembly.jl:
    1073-1111): the points at which the reference materialises a function of x (src/coordfunctions.jl:186-213)."""
    g, _ = gauss01(P)
    coords = np.asarray(coords, dtype=np.float64)
    out = np.zeros((coords.shape[0], P ** 3, 3))
    for q in range(P ** 3):
        xi = (g[q % P], g[(q // P) % P], g[q // (P * P)])
        for v in range(8):
            m = 1.0
            for d in range(3):
                m *= xi[d] if (v >> d) & 1 else 1.0 - xi[d]
            out[:, q, :] += m * coords[:, v, :]
    return out


def lagrange1d(Q, xi):
    N, dN = np.zeros(Q), np.zeros(Q)
    load().orc_lagrange1d(Q, float(xi), _p(N, C.c_double), _p(dN, C.c_double))
    return N, dN


def law_P(law, lam, mu, g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    R = np.zeros(9)
    load().orc_law_P(LAWS[law], lam, mu, _p(g, C.c_double), _p(R, C.c_double))
    return R


def law_T(law, lam, mu, g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    T = np.zeros(81)
    load().orc_law_T(LAWS[law], lam, mu, _p(g, C.c_double), _p(T, C.c_double))
    return T


def law_W(law, lam, mu, g):
    g = np.ascontiguousarray(g, dtype=np.float64)
    return load().orc_law_W(LAWS[law], lam, mu, _p(g, C.c_double))


class OracleOp:
    """The oracle's operator on the same arrays the C ABI takes (mesh.HexSpace conventions)."""

    def __init__(self, space, law="svk", lam=2.0, mu=1.0, points1d=None, dvals=None, body=None):
        L = load()
        self.space = space
        self.P = points1d or (space.order + 1)
        self.np = self.P ** 3
        self.nd = space.nd
        self.ncells = space.ncells
        self.nfree = space.nfree
        self._keep = [np.ascontiguousarray(space.cell_dofs, dtype=np.int32),
                      np.ascontiguousarray(space.nodeperm, dtype=np.int32),
                      np.ascontiguousarray(space.ldof.reshape(-1), dtype=np.int32),
                      np.ascontiguousarray(space.coords, dtype=np.float64)]
        b = None if body is None else np.ascontiguousarray(body, dtype=np.float64)
        self.h = L.orc_op_create(space.Q, self.P, space.ncells, _p(self._keep[0], C.c_int32),
                                 _p(self._keep[1], C.c_int32), _p(self._keep[2], C.c_int32),
                                 _p(self._keep[3], C.c_double), space.nfree, space.ndir, LAWS[law], lam, mu,
                                 None if b is None else _p(b, C.c_double))
        if not self.h:
            raise RuntimeError(L.orc_last_error().decode())
        if dvals is not None:
            self.set_dirichlet(dvals)

    def __del__(self):
        try:
            load().orc_op_destroy(self.h)
        except Exception:
            pass

    def set_dirichlet(self, dvals):
        self._dv = np.ascontiguousarray(dvals, dtype=np.float64)
        load().orc_op_set_dirichlet(self.h, _p(self._dv, C.c_double))

    def set_mu_field(self, mu):
        self._mu = None if mu is None else np.ascontiguousarray(mu, dtype=np.float64)
        load().orc_op_set_mu_field(self.h, None if mu is None else _p(self._mu, C.c_double))

    def colors(self):
        c = np.zeros(self.ncells, dtype=np.int32)
        load().orc_op_colors(self.h, _p(c, C.c_int32))
        return c

    def prepare(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        T = np.zeros((self.ncells, self.np, 81))
        iJt = np.zeros((self.ncells, self.np, 9))
        dV = np.zeros((self.ncells, self.np))
        load().orc_prepare(self.h, _p(x, C.c_double), _p(T, C.c_double), _p(iJt, C.c_double), _p(dV, C.c_double))
        return T, iJt, dV

    def jv(self, prep, dx, nthreads=0):
        T, iJt, dV = prep
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        y = np.zeros(self.nfree)
        load().orc_jv(self.h, _p(T, C.c_double), _p(iJt, C.c_double), _p(dV, C.c_double), _p(dx, C.c_double),
                      _p(y, C.c_double), int(nthreads))
        return y

    def energy(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return load().orc_energy(self.h, _p(x, C.c_double))

    def face_load(self, face_cell, face_local, traction):
        """The load vector of a dead traction on faces (nfaces x 3): (f on the free ids, f on the Dirichlet ids)."""
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        fl = np.ascontiguousarray(face_local, dtype=np.int32)
        tr = np.ascontiguousarray(traction, dtype=np.float64).reshape(-1, 3)
        f = np.zeros(self.nfree + self.space.ndir)
        load().orc_face_load(self.h, fc.size, _p(fc, C.c_int32), _p(fl, C.c_int32), _p(tr, C.c_double),
                             _p(f, C.c_double))
        return f[:self.nfree], f[self.nfree:]

    def face_mass(self, face_cell, face_local, stiffness):
        """The Robin term k int u.v dGamma on faces as a dense matrix over all dofs, (nfree + ndir)^2: free ids first,
        then the Dirichlet ids."""
        fc = np.ascontiguousarray(face_cell, dtype=np.int32)
        fl = np.ascontiguousarray(face_local, dtype=np.int32)
        k = np.ascontiguousarray(np.broadcast_to(np.asarray(stiffness, dtype=np.float64), (fc.size,)))
        n = self.nfree + self.space.ndir
        M = np.zeros((n, n), order="F")
        load().orc_face_mass(self.h, fc.size, _p(fc, C.c_int32), _p(fl, C.c_int32), _p(k, C.c_double),
                             M.ctypes.data_as(C.POINTER(C.c_double)))
        return np.ascontiguousarray(M)

    def point_gradients(self, x):
        """(ncells, np, 3, 3) with [cell, q, c, J] = d u_c / d x_J at every quadrature point (Dirichlet values of
        the operator)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        G = np.zeros((self.ncells, self.np, 9))
        load().orc_point_gradients(self.h, _p(x, C.c_double), _p(G, C.c_double))
        return G.reshape(self.ncells, self.np, 3, 3)  # stored [J + 3c]: row c, column J

    def residual(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        b = np.zeros(self.nfree)
        load().orc_residual(self.h, _p(x, C.c_double), _p(b, C.c_double))
        return b

    def element_jacobians(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        K = np.zeros((self.ncells, self.nd, self.nd))  # K[cell, jj, ii] = Ke[ii, jj] (column-major ii + nd*jj)
        load().orc_element_jacobians(self.h, _p(x, C.c_double), _p(K, C.c_double))
        return np.ascontiguousarray(K.transpose(0, 2, 1))  # Ke[cell, ii, jj]

    def pattern(self):
        L = load()
        nnz = L.orc_pattern_nnz(self.h)
        colptr = np.zeros(self.nfree + 1, dtype=np.int64)
        rowval = np.zeros(nnz, dtype=np.int64)
        nzpos = np.zeros((self.ncells, self.nd * self.nd), dtype=np.int32)
        L.orc_pattern(self.h, _p(colptr, C.c_int64), _p(rowval, C.c_int64), _p(nzpos, C.c_int32))
        return colptr, rowval, nzpos

    def jacobian(self, x, nzpos, nnz):
        x = np.ascontiguousarray(x, dtype=np.float64)
        nzpos = np.ascontiguousarray(nzpos, dtype=np.int32)
        nz = np.zeros(nnz)
        load().orc_jacobian(self.h, _p(x, C.c_double), _p(nzpos, C.c_int32), _p(nz, C.c_double), nnz)
        return nz

    def csc(self, x):
        """The assembled Jacobian at x as a scipy CSC matrix (pattern and values of the oracle)."""
        import scipy.sparse as sp

        colptr, rowval, nzpos = self.pattern()
        nz = self.jacobian(x, nzpos, rowval.size)
        return sp.csc_matrix((nz, rowval - 1, colptr - 1), shape=(self.nfree, self.nfree))

    def node_blocks(self):
        """field_blocks: (block_of[cell, table node], dofs[block, 3]) in first-seen order."""
        nn = self.space.nn
        block_of = np.zeros((self.ncells, nn), dtype=np.int32)
        cap = self.ncells * nn
        dofs = np.zeros((cap, 3), dtype=np.int32)
        nb = C.c_int64()
        rc = load().orc_node_blocks(self.h, _p(block_of, C.c_int32), _p(dofs, C.c_int32), cap, C.byref(nb))
        assert rc == 0
        return block_of, dofs[: nb.value].copy()

    def blocks_from_tangents(self, prep):
        """Inverted nodal blocks Binv[block, i + 3j] from the prepared tangents."""
        T, iJt, dV = prep
        block_of, dofs = self.node_blocks()
        Binv = np.zeros((dofs.shape[0], 9))
        load().orc_blocks_from_tangents(self.h, _p(T, C.c_double), _p(iJt, C.c_double), _p(dV, C.c_double),
                                        _p(block_of, C.c_int32), _p(dofs, C.c_int32), dofs.shape[0],
                                        _p(Binv, C.c_double))
        return dofs, Binv


def reference_coordinate_tangents(T, iJt, dV):
    """D[cell, q, (K,c) + 9 (M,k)] = dV sum_J sum_L iJt[J,K] T[(J,c),(L,k)] iJt[L,M] from the reference-layout
    T, iJt, dV (the storage libhfk keeps; see hfk_kernels.hip)."""
    nc, npts = dV.shape
    T4 = T.reshape(nc, npts, 3, 3, 3, 3, order="C")  # index order reversed: [k, L, c, J] for T[J + 3c + 9(L + 3k)]
    Tm = np.transpose(T4, (0, 1, 5, 4, 3, 2))  # [J, c, L, k]
    B = iJt.reshape(nc, npts, 3, 3)  # iJt[J + 3K] -> reversed [K, J]
    Bm = np.transpose(B, (0, 1, 3, 2))  # [J, K]
    D = np.einsum("xqJK,xqJcLk,xqLM->xqKcMk", Bm, Tm, Bm) * dV[:, :, None, None, None, None]
    # flatten to i + 9 j with i = K + 3c, j = M + 3k: C-order reversed indices [k, M, c, K]
    return np.ascontiguousarray(np.transpose(D, (0, 1, 5, 4, 3, 2)).reshape(nc, npts, 81))


# ---------------------------------------------------------------------------------------------------------
# Solver-layer restatements on assembled (dense or scipy) matrices, small problems only.

def chebyshev_start(n):
    """src/smoothers.jl:68-71"""
    i = np.arange(1, n + 1, dtype=np.float64)
    v = np.sin(1.0 + 0.7 * i) + 0.3 * np.cos(2.3 * i)
    return v / np.linalg.norm(v)


def block_inverse_matrix(dofs, Binv, n):
    """The sparse matrix of ldiv!(y, blocks, r) (src/preconditioners.jl:327-342)."""
    import scipy.sparse as sp

    rows, cols, vals = [], [], []
    for k in range(dofs.shape[0]):
        for i in range(3):
            if dofs[k, i] <= 0:
                continue
            for j in range(3):
                if dofs[k, j] <= 0:
                    continue
                rows.append(dofs[k, i] - 1)
                cols.append(dofs[k, j] - 1)
                vals.append(Binv[k, i + 3 * j])
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def arnoldi_bounds(Aop, start, k):
    """eigenvalue_bounds (src/preconditioners.jl:374-389): k steps of Arnoldi (modified Gram-Schmidt) on the linear
    map Aop from start; truncation at a near breakdown; extreme real parts of the Ritz values."""
    n = start.size
    k = min(k, n)
    V = np.zeros((n, k + 1))
    H = np.zeros((k + 1, k))
    V[:, 0] = start / np.linalg.norm(start)
    steps = k
    for j in range(k):
        w = Aop(V[:, j])
        for i in range(j + 1):
            H[i, j] = V[:, i] @ w
            w = w - H[i, j] * V[:, i]
        H[j + 1, j] = np.linalg.norm(w)
        if not H[j + 1, j] > 0:
            steps = j + 1
            break
        V[:, j + 1] = w / H[j + 1, j]
    jj = steps
    for i in range(1, steps):
        if abs(H[i, i - 1]) <= 1e-8 * max(1.0, np.abs(H[:i, :i]).max()):
            jj = i
            break
    ev = np.linalg.eigvals(H[:jj, :jj]).real
    return ev.min(), ev.max()


def chebyshev_smooth(A, Minv, x, rhs, a, b, degree, zero_guess=False, want_residual=False):
    """smooth! (src/smoothers.jl:116-146) with matrices A and M^-1 (anything supporting @)."""
    th, de = (b + a) / 2, (b - a) / 2
    sg = th / de
    x = np.zeros_like(rhs) if zero_guess else x.copy()
    r = rhs.copy() if zero_guess else rhs - A @ x
    z = Minv @ r
    rho = 1 / sg
    d = z / th
    for i in range(1, degree + 1):
        x = x + d
        if i == degree and not want_residual:
            break
        r = r - A @ d
        if i == degree:
            break
        z = Minv @ r
        rn = 1 / (2 * sg - rho)
        d = rn * rho * d + (2 * rn / de) * z
        rho = rn
    return (x, r) if want_residual else x


def gmres_right(A, b, Minv, rtol, restart, maxiter, atol=0.0, flexible=False):
    """Restarted right-preconditioned GMRES (Krylov.jl's gmres!/fgmres! with restart = true, as MatrixFreeSolver
    calls them, src/solvers.jl:128-133): Arnoldi on A M^-1 from r0 = b (x0 = 0) with modified Gram-Schmidt, Givens
    rotations, converged when the residual estimate |g_{j+1}| <= atol + rtol ||b||; x += M^-1 V y (or Z y).
    Returns (x, iterations, converged)."""
    n = b.size
    bn = np.linalg.norm(b)
    tol = atol + rtol * bn
    x = np.zeros(n)
    r = b.copy()
    it = 0
    beta = bn
    if bn == 0:
        return x, 0, True
    while it < maxiter:
        V = np.zeros((restart + 1, n))
        Z = np.zeros((restart, n))
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = beta
        V[0] = r / beta
        k = 0
        done = False
        for j in range(restart):
            if it >= maxiter:
                break
            Z[j] = Minv(V[j])
            w = A @ Z[j]
            for i in range(j + 1):
                H[i, j] = V[i] @ w
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            it += 1
            k = j + 1
            for i in range(j):
                a0, a1 = H[i, j], H[i + 1, j]
                H[i, j], H[i + 1, j] = cs[i] * a0 + sn[i] * a1, -sn[i] * a0 + cs[i] * a1
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            if abs(g[j + 1]) <= tol or H[j + 1, j] == 0:
                done = True
                break
            V[j + 1] = w / np.linalg.norm(w)
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + (Z[:k].T @ y if flexible else Minv(V[:k].T @ y))
        if done:
            return x, it, True
        r = b - A @ x
        beta = np.linalg.norm(r)
        if beta <= tol:
            return x, it, True
    return x, it, np.linalg.norm(b - A @ x) <= tol


def lagrange_embedding(Qc, Qf):
    """E[tf, tc]: coarse tensor shape function tc at fine tensor node tf (equispaced nodes on [0,1]^3), from the
    oracle's 1D Lagrange basis."""
    NF, NC = Qf ** 3, Qc ** 3
    E = np.zeros((NF, NC))
    for tf in range(NF):
        fi = (tf % Qf, (tf // Qf) % Qf, tf // (Qf * Qf))
        vals = [lagrange1d(Qc, f / (Qf - 1))[0] for f in fi]
        for tc in range(NC):
            ci = (tc % Qc, (tc // Qc) % Qc, tc // (Qc * Qc))
            E[tf, tc] = vals[0][ci[0]] * vals[1][ci[1]] * vals[2][ci[2]]
    E[np.abs(E) < 1e-7] = 0.0
    E[np.abs(E - 1) < 1e-7] = 1.0
    return E


def prolongation_matrix(coarse_space, fine_space):
    """The matrix of prolongate! (src/transfers.jl:188-245): fine free dofs x coarse free dofs, insertion at shared
    fine nodes (every cell writes the same row)."""
    import scipy.sparse as sp

    Qc, Qf = coarse_space.Q, fine_space.Q
    E = lagrange_embedding(Qc, Qf)
    rows = {}
    NF, NC = Qf ** 3, Qc ** 3
    for cell in range(fine_space.ncells):
        fd, cd = fine_space.cell_dofs[cell], coarse_space.cell_dofs[cell]
        for c in range(3):
            for tf in range(NF):
                idf = fd[fine_space.ldof[c, fine_space.nodeperm[tf]]]
                if idf <= 0:
                    continue
                row = {}
                for tc in range(NC):
                    idc = cd[coarse_space.ldof[c, coarse_space.nodeperm[tc]]]
                    if E[tf, tc] != 0 and idc > 0:
                        row[idc - 1] = E[tf, tc]
                rows[idf - 1] = row
    r, cidx, v = [], [], []
    for i, row in rows.items():
        for j, e in row.items():
            r.append(i)
            cidx.append(j)
            v.append(e)
    return sp.csr_matrix((v, (r, cidx)), shape=(fine_space.nfree, coarse_space.nfree))
