"""Tetrahedral Taylor-Hood P2-P1 spaces on a Cartesian box split into 6 tetrahedra per cell, the quadrature rule,
and the synthetic fields of the Oseen workload (host-side helpers; a Julia host passes Gridap's arrays instead).

Mesh. Every box cell is cut into the 6 tetrahedra of the paths from its vertex (0,0,0) to (1,1,1) along the unit
axes in each order (all share the main diagonal; a conforming mesh, as `simplexify` of a Cartesian model,
test/common.jl:167).  Every edge joins two vertices whose offset lies in {0,1}^3, so the P2 nodes (vertices and
edge midpoints) are exactly the points of the half-step lattice: (2n+1)^3 of them.

Cell dofs (the layout of hfk_oseen_create): [c*10 + a] velocity component c at P2 node a (vertices 0..3, then the
midpoints of the edges (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)), [30 + a] pressure at vertex a.  Free ids: velocity
nodes in lattice order (3 rank + c + 1), then every pressure vertex (nu + rank + 1).  Dirichlet ids: velocity
nodes on the inlet x = x0 and on the walls (y and z faces), -(3 rank + c + 1).
"""
from __future__ import annotations

from dataclasses import dataclass
from itertools import permutations

import numpy as np

from .synthetic import normal, splitmix64, uniform

EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))


@dataclass
class TetP2P1Space:
    n: tuple
    domain: tuple
    coords: np.ndarray        # (ncells, 4, 3)
    cell_dofs: np.ndarray     # (ncells, 34) int32
    nfree: int
    nfree_u: int
    ndir: int
    node_xyz: np.ndarray      # P2 lattice nodes (x fastest)
    free_node_of: np.ndarray  # free velocity node rank -> lattice node
    dir_node_of: np.ndarray   # Dirichlet velocity node rank -> lattice node
    vert_xyz: np.ndarray      # pressure vertices (x fastest)

    @property
    def ncells(self) -> int:
        return self.coords.shape[0]

    def dirichlet_values(self, fun) -> np.ndarray:
        """Dirichlet values of the velocity ``fun`` (callable (m, 3) -> (m, 3)) in the Dirichlet numbering."""
        v = np.asarray(fun(self.node_xyz[self.dir_node_of]), dtype=np.float64)
        return np.ascontiguousarray(v.reshape(-1))

    def volumes(self) -> np.ndarray:
        e = self.coords[:, 1:, :] - self.coords[:, :1, :]
        return np.abs(np.linalg.det(e)) / 6.0

    def quadrature_points(self, qpts) -> np.ndarray:
        """Physical coordinates (ncells, np, 3) of reference points qpts (np, 3) through the affine maps."""
        x0 = self.coords[:, 0, :]
        J = self.coords[:, 1:, :] - x0[:, None, :]  # rows: x_{d+1} - x0
        return x0[:, None, :] + np.einsum("qd,cdi->cqi", np.asarray(qpts), J)


def reorder_local(space: TetP2P1Space, nodeperm, pnodeperm, ldof) -> TetP2P1Space:
    """The same space with another local order, as a host whose tables differ from the library's hands it over:
    ``nodeperm[f]`` the host P2 table node of the library's node f, ``pnodeperm[v]`` the host P1 node of vertex v,
    ``ldof[c*10 + n]`` / ``ldof[30 + n]`` the cell-local dof of (component c, host P2 node n) / (pressure, host P1
    node n).  The cell dofs are rewritten in that order and the three tables attached (hfk_oseen_create reads
    them)."""
    from copy import copy

    nodeperm, pnodeperm, ldof = (np.asarray(a, dtype=np.int32) for a in (nodeperm, pnodeperm, ldof))
    perm = np.empty(34, dtype=np.int64)  # perm[i] = host local dof of the library's local dof i
    for c in range(3):
        perm[c * 10:(c + 1) * 10] = ldof[c * 10 + nodeperm]
    perm[30:] = ldof[30 + pnodeperm]
    out = copy(space)
    cd = np.empty_like(space.cell_dofs)
    cd[:, perm] = space.cell_dofs
    out.cell_dofs = cd
    out.nodeperm, out.pnodeperm, out.ldof, out.local_perm = nodeperm, pnodeperm, ldof, perm
    return out


def tet_p2p1_space(n, domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)) -> TetP2P1Space:
    nx, ny, nz = (int(v) for v in n)
    x0, x1, y0, y1, z0, z1 = domain
    h = np.array([(x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz])
    Lx, Ly, Lz = 2 * nx + 1, 2 * ny + 1, 2 * nz + 1
    Vx, Vy = nx + 1, ny + 1
    # tetrahedra: 6 per box cell, box cells x fastest
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = (a.transpose(2, 1, 0).reshape(-1) for a in (ex, ey, ez))
    base = np.stack([ex, ey, ez], axis=1)  # (nbox, 3) integer vertex coordinates
    perms = list(permutations(range(3)))
    nbox = base.shape[0]
    V = np.zeros((nbox, 6, 4, 3), dtype=np.int64)  # integer vertex coords of each tet
    for t, pm in enumerate(perms):
        p = np.zeros(3, dtype=np.int64)
        V[:, t, 0] = base
        for k in range(3):
            p[pm[k]] += 1
            V[:, t, k + 1] = base + p
    V = V.reshape(-1, 4, 3)
    ncells = V.shape[0]
    coords = np.array([x0, y0, z0])[None, None, :] + V * h[None, None, :]
    # P2 nodes on the doubled lattice: vertex 2v, edge midpoint v_a + v_b
    D = np.zeros((ncells, 10, 3), dtype=np.int64)
    D[:, :4] = 2 * V
    for e, (a, b) in enumerate(EDGES):
        D[:, 4 + e] = V[:, a] + V[:, b]
    gnode = D[..., 0] + Lx * (D[..., 1] + Ly * D[..., 2])  # (ncells, 10)
    gvert = V[..., 0] + Vx * (V[..., 1] + Vy * V[..., 2])  # (ncells, 4)
    # lattice nodes and Dirichlet marks (inlet x = x0 and the four walls)
    I, J, K = np.meshgrid(np.arange(Lx), np.arange(Ly), np.arange(Lz), indexing="ij")
    I, J, K = (a.transpose(2, 1, 0).reshape(-1) for a in (I, J, K))
    node_xyz = np.stack([x0 + I * h[0] / 2, y0 + J * h[1] / 2, z0 + K * h[2] / 2], axis=1)
    isdir = (I == 0) | (J == 0) | (J == Ly - 1) | (K == 0) | (K == Lz - 1)
    free_rank = np.cumsum(~isdir) - 1
    dir_rank = np.cumsum(isdir) - 1
    nfree_nodes, ndir_nodes = int((~isdir).sum()), int(isdir.sum())
    nfree_u = 3 * nfree_nodes
    nvert = (nx + 1) * (ny + 1) * (nz + 1)
    cd = np.zeros((ncells, 34), dtype=np.int64)
    for c in range(3):
        idn = np.where(isdir[gnode], -(3 * dir_rank[gnode] + c + 1), 3 * free_rank[gnode] + c + 1)
        cd[:, c * 10:(c + 1) * 10] = idn
    cd[:, 30:34] = nfree_u + gvert + 1
    VI, VJ, VK = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    VI, VJ, VK = (a.transpose(2, 1, 0).reshape(-1) for a in (VI, VJ, VK))
    vert_xyz = np.stack([x0 + VI * h[0], y0 + VJ * h[1], z0 + VK * h[2]], axis=1)
    return TetP2P1Space(n=(nx, ny, nz), domain=tuple(domain), coords=coords,
                        cell_dofs=np.ascontiguousarray(cd.astype(np.int32)), nfree=nfree_u + nvert,
                        nfree_u=nfree_u, ndir=3 * ndir_nodes, node_xyz=node_xyz,
                        free_node_of=np.nonzero(~isdir)[0], dir_node_of=np.nonzero(isdir)[0], vert_xyz=vert_xyz)


def tet_quadrature_deg5():
    """The symmetric 14-point rule on the unit tetrahedron (positive weights, exact for total degree 5): points
    (np, 3) in reference coordinates and weights (np,) summing to 1/6."""
    a, b = 0.0927352503108912, 0.3108859192633006
    c = 0.0455037041256496
    wa, wb, wc = 0.01224884051939366, 0.01878132095300264, 0.007091003462846911
    bary = []
    wts = []
    for t, w in ((a, wa), (b, wb)):
        for k in range(4):
            l4 = [t] * 4
            l4[k] = 1.0 - 3.0 * t
            bary.append(l4)
            wts.append(w)
    d = 0.5 - c
    for (i, j) in EDGES:  # two barycentrics equal c, the other two equal 0.5 - c
        l4 = [d] * 4
        l4[i] = c
        l4[j] = c
        bary.append(l4)
        wts.append(wc)
    bary = np.array(bary)
    return np.ascontiguousarray(bary[:, 1:]), np.array(wts)


def tri_quadrature(degree: int):
    """Collapsed Gauss rule on the reference triangle {s, t >= 0, s + t <= 1}: Gauss-Legendre in a and Gauss-Jacobi
    with weight (1 - b) in b, s = a (1 - b), t = b; n = floor(degree / 2) + 1 points per direction, exact for total
    degree <= 2n - 1.  Points (n^2, 2) and weights summing to 1/2."""
    from scipy.special import roots_jacobi

    n = degree // 2 + 1
    ga, wa = np.polynomial.legendre.leggauss(n)
    a, wa = (ga + 1.0) / 2.0, wa / 2.0
    tb, wb = roots_jacobi(n, 1.0, 0.0)
    bb, wb = (tb + 1.0) / 2.0, wb / 4.0
    pts = np.array([(ai * (1.0 - bj), bj) for bj in bb for ai in a])
    wts = np.array([wai * wbj for wbj in wb for wai in wa])
    return pts, wts


REF_VERTS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


@dataclass
class OutletFaces:
    """Boundary faces of one plane with their quadratures, in the layout of hfk_oseen_set_outlet.  ``xo``/``xb``:
    (nfaces, nq, 4) = reference coordinates of the point in the face's cell and the weight dS; ``xb_phys``: the
    physical coordinates of the dGb points (where the host evaluates w)."""
    cell: np.ndarray
    normal: np.ndarray
    xo: np.ndarray
    xb: np.ndarray
    xb_phys: np.ndarray
    area: float


def _face_points(space, cell, fv, tri_pts, tri_w):
    """Reference coordinates in the cell, physical coordinates and weights dS of the triangle rule mapped onto the
    faces with local vertices fv (nfaces, 3)."""
    R = REF_VERTS[fv]  # (nf, 3, 3)
    xi = R[:, None, 0, :] + tri_pts[None, :, 0, None] * (R[:, None, 1, :] - R[:, None, 0, :]) + \
        tri_pts[None, :, 1, None] * (R[:, None, 2, :] - R[:, None, 0, :])
    X = space.coords[cell]  # (nf, 4, 3)
    P = X[np.arange(len(cell))[:, None], fv]  # face vertices (nf, 3, 3)
    area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    J = X[:, 1:, :] - X[:, :1, :]
    phys = X[:, None, 0, :] + np.einsum("fqd,fdi->fqi", xi, J)
    dS = 2.0 * area[:, None] * tri_w[None, :]
    return xi, phys, dS, area


def outlet_faces(space, degree_o: int = 4, degree_b: int = 5, axis: int = 0, side: int = 1) -> OutletFaces:
    """The faces of the tetrahedra on the box plane x_axis = (domain max if side else min), with the triangle rules
    of degree degree_o (Robin and pressure terms, Measure(Γo, 4)) and degree_b (backflow, Measure(Γo, 5)),
    test/common.jl:197."""
    plane = space.domain[2 * axis + (1 if side else 0)]
    h = (space.domain[2 * axis + 1] - space.domain[2 * axis]) / space.n[axis]
    on = np.abs(space.coords[:, :, axis] - plane) < 1e-9 * max(1.0, abs(plane), h)
    cell = np.nonzero(on.sum(axis=1) == 3)[0]
    fv = np.array([np.nonzero(r)[0] for r in on[cell]], dtype=np.int64).reshape(-1, 3)
    opp = np.array([np.nonzero(~r)[0][0] for r in on[cell]], dtype=np.int64)
    X = space.coords[cell]
    P = X[np.arange(len(cell))[:, None], fv]
    nrm = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    inward = np.einsum("fi,fi->f", X[np.arange(len(cell)), opp] - P[:, 0], nrm)
    nrm *= np.where(inward > 0, -1.0, 1.0)[:, None]
    po, wo = tri_quadrature(degree_o)
    pb, wb = tri_quadrature(degree_b)
    xio, _, dSo, area = _face_points(space, cell, fv, po, wo)
    xib, physb, dSb, _ = _face_points(space, cell, fv, pb, wb)
    xo = np.concatenate([xio, dSo[..., None]], axis=2)
    xb = np.concatenate([xib, dSb[..., None]], axis=2)
    return OutletFaces(cell=cell.astype(np.int32), normal=np.ascontiguousarray(nrm), xo=np.ascontiguousarray(xo),
                       xb=np.ascontiguousarray(xb), xb_phys=physb, area=float(area.sum()))


def element_velocity(space, x, dvals) -> np.ndarray:
    """Nodal velocity values (ncells, 3, 10) of the cells from the free values x and the Dirichlet values."""
    ids = np.asarray(space.cell_dofs)[:, :30]
    x, dvals = np.asarray(x), np.asarray(dvals)
    ue = np.where(ids > 0, x[np.maximum(ids, 1) - 1], dvals[np.maximum(-ids, 1) - 1])
    return ue.reshape(-1, 3, 10)


def p2_laplacian(space, x, dvals) -> np.ndarray:
    """Elementwise Laplacian (ncells, 3) of the P2 velocity with free values x and Dirichlet values dvals: constant
    on each affine cell (Laplacians of the shape functions: 4 |grad l_a|^2 at vertex a, 8 grad l_a . grad l_b at the
    midpoint of edge (a, b)).  The reference obtains the same field by interpolating grad(un) . e_k into a
    discontinuous P1 space, which holds the (linear) gradient of a P2 field exactly, and taking its divergence
    (test/common.jl:191-195)."""
    GL = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    J = np.transpose(space.coords[:, 1:, :] - space.coords[:, :1, :], (0, 2, 1))  # J[c][i][d]
    G = np.einsum("kd,cdi->cki", GL, np.linalg.inv(J))  # physical gradients of the barycentrics (ncells, 4, 3)
    lapN = np.zeros((space.ncells, 10))
    for a in range(4):
        lapN[:, a] = 4.0 * np.einsum("ci,ci->c", G[:, a], G[:, a])
    for e, (a, b) in enumerate(EDGES):
        lapN[:, 4 + e] = 8.0 * np.einsum("ci,ci->c", G[:, a], G[:, b])
    return np.einsum("cka,ca->ck", element_velocity(space, x, dvals), lapN)


def divfree_field(seed: int, modes: int = 12, kmax: int = 2):
    """w(x) = sum_m (a_m x k_m/|k_m|) cos(2 pi k_m . x + phi_m): k_m uniform in {-kmax..kmax}^3 without 0,
    a_m ~ N(0, I), phi_m ~ U[0, 2 pi), all from SplitMix64(seed); divergence free term by term.  Returns the
    callable (m, 3) -> (m, 3) (unscaled)."""
    ks = []
    r = splitmix64(seed, 64 * modes)
    i = 0
    while len(ks) < modes:
        k = (r[3 * i: 3 * i + 3] % np.uint64(2 * kmax + 1)).astype(np.int64) - kmax
        i += 1
        if np.any(k != 0):
            ks.append(k.astype(np.float64))
    K = np.array(ks)
    A = normal(seed + 1, 3 * modes).reshape(modes, 3)
    phi = uniform(seed + 2, modes) * 2 * np.pi
    khat = K / np.linalg.norm(K, axis=1, keepdims=True)
    amp = np.cross(A, khat)  # (modes, 3), orthogonal to k

    def w(xyz):
        xyz = np.asarray(xyz, dtype=np.float64)
        arg = 2 * np.pi * xyz @ K.T + phi[None, :]
        return np.cos(arg) @ amp

    return w
