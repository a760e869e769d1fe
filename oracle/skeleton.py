"""Face integrals of the hexahedral operator -- skeleton (interior facet) integrals and boundary groups with a
captured integrand -- restated in numpy for small meshes.

Test infrastructure: only tests/ and the checker import this module; the product never does.

It follows the reference's skeleton path:
* src/skeleton.jl:1-16 -- a facet couples its plus and minus cell; the facet quadrature points are mapped into the
  reference cell of each side, the shape functions of each side are tabulated there, the measure is the plus face's
  (``w meas(Jf+ . Jt+)``), the point form sees the unknown of both sides;
* src/skeleton.jl:291-330 (``cells!``) -- per facet: gather both cells, element residual / Jacobian over the points,
  scatter both halves;
* src/skeleton.jl:39-69 (``color_facets``) and :355-390 (``skeleton_keys!``: the cross blocks of every facet);
* src/assembly.jl:539-578 -- a boundary group: the same with one cell per face.

A point form here is an object with ``residual(s, x, n, h, p)``, ``jvp(s, ds, x, n, h, p)`` and
``energy(s, x, n, h, p)`` of the 12 slots per side s (plus: u 0..2, du_c/dx_K at 3 + 3c + K; minus: the same at
12..23), the physical point x, the unit normal n outward from the plus cell, h = sqrt(face area) and the parameters
p -- the interface of the device forms.  Faces: a skeleton's (plus cell, plus face, minus cell, minus face,
orientation) arrays, or a boundary group's (cell, face) arrays.
"""
from __future__ import annotations

import numpy as np

from oracle.oracle import gauss01, lagrange1d


def face_dirs(lf):
    d, side = lf >> 1, lf & 1
    e1 = 1 if d == 0 else 0
    e2 = 1 if d == 2 else 2
    return d, side, e1, e2


def face_xi(lf, s, t):
    d, side, e1, e2 = face_dirs(lf)
    xi = np.zeros(3)
    xi[d], xi[e1], xi[e2] = side, s, t
    return xi


def orient_st(o, s, t):
    s2, t2 = (t, s) if o & 1 else (s, t)
    if o & 2:
        s2 = 1.0 - s2
    if o & 4:
        t2 = 1.0 - t2
    return s2, t2


def trilinear_jacobian(X, xi):
    """J[i, K] = dx_i/dxi_K of the trilinear map of the vertices X (8, 3), vertex v = bx + 2 by + 4 bz."""
    J = np.zeros((3, 3))
    for v in range(8):
        b = [(v >> k) & 1 for k in range(3)]
        m = [xi[k] if b[k] else 1.0 - xi[k] for k in range(3)]
        dm = [1.0 if b[k] else -1.0 for k in range(3)]
        g = np.array([dm[0] * m[1] * m[2], m[0] * dm[1] * m[2], m[0] * m[1] * dm[2]])
        J += np.outer(X[v], g)
    return J


def shape(Q, xi):
    """Values (NN,) and reference gradients (NN, 3) of the tensor Lagrange basis, node t = i + Q j + Q^2 k."""
    n0, d0 = lagrange1d(Q, xi[0])
    n1, d1 = lagrange1d(Q, xi[1])
    n2, d2 = lagrange1d(Q, xi[2])
    NN = Q ** 3
    N, G = np.zeros(NN), np.zeros((NN, 3))
    for t in range(NN):
        i, j, k = t % Q, (t // Q) % Q, t // (Q * Q)
        N[t] = n0[i] * n1[j] * n2[k]
        G[t] = (d0[i] * n1[j] * n2[k], n0[i] * d1[j] * n2[k], n0[i] * n1[j] * d2[k])
    return N, G


def trilinear_point(X, xi):
    return sum(((xi[0] if v & 1 else 1 - xi[0]) * (xi[1] if v & 2 else 1 - xi[1]) * (xi[2] if v & 4 else 1 - xi[2])) * X[v]
               for v in range(8))


def facet_tables(space, P, plus_cell, plus_face, minus_cell=None, minus_face=None, orient=0):
    """Per point q = a + P b of the face: (dS, n, x, [per side (shape values (NN,), physical gradients (NN, 3))]);
    one side without a minus cell."""
    Q = space.Q
    g, w = gauss01(P)
    cells = [(space.coords[plus_cell], plus_face, False)]
    if minus_cell is not None:
        cells.append((space.coords[minus_cell], minus_face, True))
    pts = []
    for b in range(P):
        for a in range(P):
            s, t = g[a], g[b]
            sides = []
            for X, lf, minus in cells:
                xi = face_xi(lf, *(orient_st(orient, s, t) if minus else (s, t)))
                J = trilinear_jacobian(X, xi)
                N, Gr = shape(Q, xi)
                sides.append((N, Gr @ np.linalg.inv(J), J, xi))
            d, side, e1, e2 = face_dirs(plus_face)
            J = sides[0][2]
            n = np.cross(J[:, e1], J[:, e2])
            nn = np.linalg.norm(n)
            sg = 1.0 if ((n @ J[:, d] > 0) == (side == 1)) else -1.0
            x = trilinear_point(cells[0][0], sides[0][3])
            pts.append((w[a] * w[b] * nn, sg * n / nn, x, [sd[:2] for sd in sides]))
    return pts


def tensor_ids(space, cell):
    """Signed dof ids [c][t] of a cell in tensor node order."""
    ids = space.cell_dofs[cell]
    return np.array([[ids[space.ldof[c, space.nodeperm[t]]] for t in range(space.nn)] for c in range(3)])


def _values(ids, xf, dvals):
    out = np.zeros(ids.shape)
    out[ids > 0] = xf[ids[ids > 0] - 1]
    if dvals is not None:
        out[ids < 0] = dvals[-ids[ids < 0] - 1]
    return out


def _slots(pt, Us):
    s = np.zeros(12 * len(Us))
    for side, ((N, Gx), U) in enumerate(zip(pt[3], Us)):
        for c in range(3):
            s[12 * side + c] = N @ U[c]
            s[12 * side + 3 + 3 * c: 12 * side + 6 + 3 * c] = Gx.T @ U[c]
    return s


def _contract(pt, R):
    """Element vectors of every side ([c][t] each) of the slot coefficients R (12 per side) at one point."""
    out = []
    for side, (N, Gx) in enumerate(pt[3]):
        r = R[12 * side: 12 * side + 12]
        out.append(np.array([r[c] * N + Gx @ r[3 + 3 * c: 6 + 3 * c] for c in range(3)]))
    return out


def _faces(faces):
    """(plus cell, plus face, minus cell, minus face, orientation) per face; None for the minus side of a boundary
    group's faces."""
    if len(faces) == 5:
        return [tuple(int(a[f]) for a in faces) for f in range(len(faces[0]))]
    return [(int(faces[0][f]), int(faces[1][f]), None, None, 0) for f in range(len(faces[0]))]


def face_terms(space, faces, form, params, xf, dvals, P=None, dx=None):
    """Residual r (nfree,), energy, and (if dx is given) the product J dx (nfree,) of the face integrals at state xf."""
    P = P or space.Q
    r = np.zeros(space.nfree)
    y = np.zeros(space.nfree)
    energy = 0.0
    for pc, pf, mc, mf, ori in _faces(faces):
        pts = facet_tables(space, P, pc, pf, mc, mf, ori)
        h = np.sqrt(sum(pt[0] for pt in pts))
        ids = [tensor_ids(space, c) for c in ((pc,) if mc is None else (pc, mc))]
        Us = [_values(i, xf, dvals) for i in ids]
        dUs = [_values(i, dx, None) for i in ids] if dx is not None else None
        for pt in pts:
            s = _slots(pt, Us)
            energy += pt[0] * form.energy(s, pt[2], pt[1], h, params)
            pairs = [(r, form.residual(s, pt[2], pt[1], h, params))]
            if dx is not None:
                pairs.append((y, form.jvp(s, _slots(pt, dUs), pt[2], pt[1], h, params)))
            for vec, R in pairs:
                for i, e in zip(ids, _contract(pt, pt[0] * R)):
                    m = i > 0
                    np.add.at(vec, i[m] - 1, e[m])
    return r, energy, (y if dx is not None else None)


def face_matrix(space, faces, form, params, xf, dvals, P=None):
    """The dense Jacobian (nfree, nfree) of the face integrals at state xf: columns of the form's jvp in the unit
    directions."""
    P = P or space.Q
    A = np.zeros((space.nfree, space.nfree))
    nd = space.nd
    for pc, pf, mc, mf, ori in _faces(faces):
        pts = facet_tables(space, P, pc, pf, mc, mf, ori)
        h = np.sqrt(sum(pt[0] for pt in pts))
        cells = (pc,) if mc is None else (pc, mc)
        ns = len(cells)
        tids = [tensor_ids(space, c) for c in cells]
        Us = [_values(i, xf, dvals) for i in tids]
        ids = np.concatenate([i.reshape(-1) for i in tids])
        K = np.zeros((ns * nd, ns * nd))
        for pt in pts:
            s = _slots(pt, Us)
            for j in range(ns * nd):
                if ids[j] <= 0:
                    continue
                e = np.zeros(ns * nd)
                e[j] = 1.0
                ds = _slots(pt, [e[k * nd:(k + 1) * nd].reshape(3, -1) for k in range(ns)])
                K[:, j] += np.concatenate([v.reshape(-1) for v in _contract(pt, pt[0] * form.jvp(s, ds, pt[2], pt[1], h, params))])
        m = ids > 0
        g = ids[m] - 1  # the face nodes appear on both sides: accumulate, do not overwrite
        np.add.at(A, (g[:, None], g[None, :]), K[np.ix_(m, m)])
    return A


skeleton_terms = face_terms
skeleton_matrix = face_matrix


def skeleton_pairs(space, facets):
    """The (column, row) pairs (1-based free ids) of the cross blocks of every facet (skeleton_keys! with no mask)."""
    pc, _, mc, _, _ = facets
    pairs = set()
    for f in range(len(pc)):
        a = [int(i) for i in space.cell_dofs[pc[f]] if i > 0]
        b = [int(i) for i in space.cell_dofs[mc[f]] if i > 0]
        pairs.update((c, r) for c in b for r in a)
        pairs.update((c, r) for c in a for r in b)
    return pairs


class GradJump:
    """The built-in facet form: energy (gamma h / 2) |grad u+ - grad u-|^2, gamma = p[0]."""

    @staticmethod
    def residual(s, x, n, h, p):
        R = np.zeros(24)
        j = s[3:12] - s[15:24]
        R[3:12] = p[0] * h * j
        R[15:24] = -p[0] * h * j
        return R

    @staticmethod
    def jvp(s, ds, x, n, h, p):
        return GradJump.residual(ds, x, n, h, p)

    @staticmethod
    def energy(s, x, n, h, p):
        j = s[3:12] - s[15:24]
        return 0.5 * p[0] * h * (j @ j)


class NonlinearFacet:
    """A nonlinear facet form with value slots, for the compiled-form path:
    e = gamma h (|J|^2 / 2 + kappa |J|^4 / 4) + beta |J n|^2 / 2 + alpha |um|^2 |J|^2 / 2, with J = grad u+ - grad u-
    (3x3, J[c, K]), um = (u+ + u-) / 2 and p = (gamma, kappa, beta, alpha)."""

    @staticmethod
    def _parts(s, n):
        J = (s[3:12] - s[15:24]).reshape(3, 3)
        um = 0.5 * (s[0:3] + s[12:15])
        return J, um

    @staticmethod
    def energy(s, x, n, h, p):
        g, k, b, a = p[:4]
        J, um = NonlinearFacet._parts(s, n)
        jj = np.sum(J * J)
        jn = J @ n
        return g * h * (0.5 * jj + 0.25 * k * jj * jj) + 0.5 * b * (jn @ jn) + 0.5 * a * (um @ um) * jj

    @staticmethod
    def residual(s, x, n, h, p):
        g, k, b, a = p[:4]
        J, um = NonlinearFacet._parts(s, n)
        jj = np.sum(J * J)
        jn = J @ n
        dJ = g * h * (1.0 + k * jj) * J + b * np.outer(jn, n) + a * (um @ um) * J  # de/dJ
        dum = a * jj * um  # de/dum
        R = np.zeros(24)
        R[3:12] = dJ.reshape(-1)
        R[15:24] = -dJ.reshape(-1)
        R[0:3] = 0.5 * dum
        R[12:15] = 0.5 * dum
        return R

    @staticmethod
    def jvp(s, ds, x, n, h, p):
        g, k, b, a = p[:4]
        J, um = NonlinearFacet._parts(s, n)
        dJ, dum = NonlinearFacet._parts(ds, n)
        jj = np.sum(J * J)
        djj = 2.0 * np.sum(J * dJ)
        dn = dJ @ n
        umum = um @ um
        dumum = 2.0 * (um @ dum)
        ddJ = (g * h * ((1.0 + k * jj) * dJ + k * djj * J) + b * np.outer(dn, n) + a * (umum * dJ + dumum * J))
        ddum = a * (djj * um + jj * dum)
        R = np.zeros(24)
        R[3:12] = ddJ.reshape(-1)
        R[15:24] = -ddJ.reshape(-1)
        R[0:3] = 0.5 * ddum
        R[12:15] = 0.5 * ddum
        return R


class FollowerPressure:
    """The built-in boundary form: follower pressure p[0], R_c = p (cof F N)_c with F = I + grad u (the virtual work
    of the nominal traction -p cof(F) N); no energy."""

    @staticmethod
    def _F(s):
        return np.eye(3) + s[3:12].reshape(3, 3)

    @staticmethod
    def residual(s, x, n, h, p):
        F = FollowerPressure._F(s)
        cof = np.linalg.det(F) * np.linalg.inv(F).T
        R = np.zeros(12)
        R[0:3] = p[0] * cof @ n
        return R

    @staticmethod
    def jvp(s, ds, x, n, h, p):
        F = FollowerPressure._F(s)
        dF = ds[3:12].reshape(3, 3)
        # d cof(F) = cof(F) tr(F^-1 dF) - cof(F) dF^T F^-T  (from cof F = det(F) F^-T)
        Fi = np.linalg.inv(F)
        cof = np.linalg.det(F) * Fi.T
        dcof = cof * np.trace(Fi @ dF) - cof @ dF.T @ Fi.T
        R = np.zeros(12)
        R[0:3] = p[0] * dcof @ n
        return R

    @staticmethod
    def energy(s, x, n, h, p):
        return 0.0


class NonlinearBoundary:
    """A nonlinear boundary form with every slot and the point: e = k |u|^2 / 2 + kappa |u|^4 / 4 + beta |grad u n|^2 / 2
    + gamma x_0 |u|^2 / 2, p = (k, kappa, beta, gamma)."""

    @staticmethod
    def energy(s, x, n, h, p):
        k, kap, b, g = p[:4]
        u = s[0:3]
        G = s[3:12].reshape(3, 3)
        uu = u @ u
        gn = G @ n
        return 0.5 * k * uu + 0.25 * kap * uu * uu + 0.5 * b * (gn @ gn) + 0.5 * g * x[0] * uu

    @staticmethod
    def residual(s, x, n, h, p):
        k, kap, b, g = p[:4]
        u = s[0:3]
        G = s[3:12].reshape(3, 3)
        uu = u @ u
        R = np.zeros(12)
        R[0:3] = (k + kap * uu + g * x[0]) * u
        R[3:12] = b * np.outer(G @ n, n).reshape(-1)
        return R

    @staticmethod
    def jvp(s, ds, x, n, h, p):
        k, kap, b, g = p[:4]
        u, du = s[0:3], ds[0:3]
        dG = ds[3:12].reshape(3, 3)
        uu = u @ u
        R = np.zeros(12)
        R[0:3] = (k + kap * uu + g * x[0]) * du + 2.0 * kap * (u @ du) * u
        R[3:12] = b * np.outer(dG @ n, n).reshape(-1)
        return R
