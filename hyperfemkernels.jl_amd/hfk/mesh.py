"""Cartesian hexahedral meshes and nodal vector Lagrangian (Q_k) spaces on them.

This is the host-side stand-in for what Gridap owns in the reference (mesh, DOF numbering, Dirichlet split,
reference-element node order).  The Julia host passes Gridap's own ``get_cell_dof_ids`` through the same
C-ABI arguments, so the kernels never depend on the numbering built here; these arrays only have to follow the
same conventions:

* ``cell_dofs[cell, j]`` -- Gridap's signed ids of the local dof ``j`` (``> 0`` free, 1-based; ``< 0``
  Dirichlet, ``-1``-based), as ``cell_dof_matrix`` (src/tables.jl:329-342) stores them;
* ``ldof[c, n]`` -- local dof of table node ``n`` and component ``c`` (``FieldTable.ldof``,
  src/tables.jl:171-237), component-major as Gridap's vector Lagrangian elements (``CONTIG``);
* ``nodeperm[t]`` -- table node of the tensor node ``t = i + Q j + Q^2 k`` (``TPTable.nodeperm``,
  src/tensorproduct.jl:13-91);
* ``coords[cell, v, :]`` -- the 8 vertices of the (trilinear) cell map, tensor order ``v = i + 2 j + 4 k``.

Cells are numbered lexicographically (x fastest) as in Gridap's ``CartesianDiscreteModel``.  The table node
order (vertices, edges, faces, interior) and the global numbering (lattice order, components interleaved) are
stand-ins: parity with Gridap's own numbering is by construction only when the Julia host passes Gridap's
arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def table_node_order(Q: int) -> np.ndarray:
    """Tensor nodes ``t`` listed in table order: vertices, then edge, face and interior nodes.

    Returns ``order`` with ``order[n] = t``; ``nodeperm`` is its inverse.  Within one class nodes are sorted
    by the directions they are interior in, then lexicographically (z slowest).
    """
    ts = []
    for t in range(Q ** 3):
        i, j, k = t % Q, (t // Q) % Q, t // (Q * Q)
        inner = tuple(int(0 < v < Q - 1) for v in (i, j, k))
        ts.append((sum(inner), tuple(-x for x in inner), k, j, i, t))
    ts.sort()
    return np.array([e[-1] for e in ts], dtype=np.int32)


@dataclass
class HexSpace:
    """A vector (3-component) Q_k space on a Cartesian box mesh, with its Dirichlet split."""

    n: tuple
    order: int
    domain: tuple
    cell_dofs: np.ndarray  # (ncells, nd) int32, table order
    ldof: np.ndarray  # (3, nn) int32
    nodeperm: np.ndarray  # (nn,) int32
    coords: np.ndarray  # (ncells, 8, 3) float64
    nfree: int
    ndir: int
    node_xyz: np.ndarray  # (nnodes_lattice, 3)
    node_free: np.ndarray  # (nnodes_lattice,) bool: all components free (Dirichlet masks are per node)
    free_node_of: np.ndarray  # free id (0-based)//3 -> lattice node
    dir_node_of: np.ndarray  # Dirichlet id (0-based)//3 -> lattice node
    lattice: tuple = field(default=())

    @property
    def Q(self) -> int:
        return self.order + 1

    @property
    def nn(self) -> int:
        return self.Q ** 3

    @property
    def nd(self) -> int:
        return 3 * self.nn

    @property
    def ncells(self) -> int:
        return self.cell_dofs.shape[0]

    def free_xyz(self) -> np.ndarray:
        """Coordinates of the node of each free dof (component order interleaved)."""
        return np.repeat(self.node_xyz[self.free_node_of], 3, axis=0)

    def dirichlet_values(self, fun) -> np.ndarray:
        """Dirichlet values of ``fun(xyz) -> (m, 3)`` at the Dirichlet dofs."""
        if self.ndir == 0:
            return np.zeros(0)
        v = np.asarray(fun(self.node_xyz[self.dir_node_of]), dtype=np.float64).reshape(-1, 3)
        return v.reshape(-1).copy()

    def lattice_dofs(self) -> np.ndarray:
        """Signed dof ids of every lattice node and component, (nnodes, 3), lattice order (x fastest)."""
        nn = self.node_xyz.shape[0]
        out = np.zeros((nn, 3), dtype=np.int32)
        for c in range(3):
            out[self.free_node_of, c] = 3 * np.arange(self.free_node_of.size) + c + 1
            out[self.dir_node_of, c] = -(3 * np.arange(self.dir_node_of.size) + c + 1)
        return out

    def quadrature_points(self, points1d: int) -> np.ndarray:
        """Physical coordinates (ncells, P^3, 3) of the tensor Gauss-Legendre points of every cell (P = points1d
        per direction, x fastest), through the trilinear map of the cell's vertices."""
        g = (np.polynomial.legendre.leggauss(int(points1d))[0] + 1.0) / 2.0
        P = g.size
        xi = np.array([[g[a], g[b], g[c]] for c in range(P) for b in range(P) for a in range(P)])
        xq = np.zeros((self.ncells, P ** 3, 3))
        for v in range(8):
            a, b, c = v & 1, (v >> 1) & 1, (v >> 2) & 1
            w = (xi[:, 0] if a else 1 - xi[:, 0]) * (xi[:, 1] if b else 1 - xi[:, 1]) * (xi[:, 2] if c else 1 - xi[:, 2])
            xq += w[None, :, None] * self.coords[:, v, None, :]
        return xq

    def free_values(self, fun) -> np.ndarray:
        v = np.asarray(fun(self.node_xyz[self.free_node_of]), dtype=np.float64).reshape(-1, 3)
        return v.reshape(-1).copy()


def dirichlet_open_x_faces(I, J, K, L):
    """Nodes in the open faces x = 0 and x = 1 of the box (Gridap's ``tag_25``/``tag_26``: face entities only,
    their edges and vertices stay free)."""
    Lx, Ly, Lz = L
    onx = (I == 0) | (I == Lx - 1)
    inner = (J > 0) & (J < Ly - 1) & (K > 0) & (K < Lz - 1)
    return onx & inner


def dirichlet_closed_x_faces(I, J, K, L):
    """Nodes on the closed faces x = 0 and x = 1 (faces with their edges and vertices: the same closed set on
    every p-multigrid level, which nested transfers require)."""
    return (I == 0) | (I == L[0] - 1)


def dirichlet_closed_x0(I, J, K, L):
    """Nodes on the closed face x = 0 (face, its edges and vertices)."""
    return I == 0


def dirichlet_none(I, J, K, L):
    return np.zeros_like(I, dtype=bool)


def hex_space(n, order: int, domain=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), dirichlet=dirichlet_open_x_faces,
              perturb: float = 0.0, seed: int = 0) -> HexSpace:
    """Vector Q_``order`` space on the Cartesian mesh ``n = (nx, ny, nz)`` of ``domain``.

    ``dirichlet(I, J, K, L)`` marks Dirichlet lattice nodes.  ``perturb`` > 0 moves interior mesh vertices by
    up to ``perturb`` times the cell size (non-affine trilinear cells, to exercise per-point geometry); the
    lattice coordinates then come from the trilinear map of each cell.
    """
    nx, ny, nz = (int(v) for v in n)
    k = int(order)
    Q = k + 1
    x0, x1, y0, y1, z0, z1 = domain
    Lx, Ly, Lz = k * nx + 1, k * ny + 1, k * nz + 1
    L = (Lx, Ly, Lz)
    nnodes = Lx * Ly * Lz
    # vertex grid (possibly perturbed)
    vx = np.linspace(x0, x1, nx + 1)
    vy = np.linspace(y0, y1, ny + 1)
    vz = np.linspace(z0, z1, nz + 1)
    VX, VY, VZ = np.meshgrid(vx, vy, vz, indexing="ij")
    verts = np.stack([VX, VY, VZ], axis=-1)  # (nx+1, ny+1, nz+1, 3)
    if perturb > 0:
        rng = np.random.default_rng(seed)
        h = np.array([(x1 - x0) / nx, (y1 - y0) / ny, (z1 - z0) / nz])
        d = (rng.random(verts.shape) - 0.5) * 2 * perturb * h
        d[0, :, :, :] = 0
        d[-1, :, :, :] = 0
        d[:, 0, :, :] = 0
        d[:, -1, :, :] = 0
        d[:, :, 0, :] = 0
        d[:, :, -1, :] = 0
        verts = verts + d
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    # cells lexicographic, x fastest
    ex = ex.transpose(2, 1, 0).reshape(-1)
    ey = ey.transpose(2, 1, 0).reshape(-1)
    ez = ez.transpose(2, 1, 0).reshape(-1)
    ncells = ex.size
    coords = np.empty((ncells, 8, 3))
    for v in range(8):
        a, b, c = v & 1, (v >> 1) & 1, (v >> 2) & 1
        coords[:, v, :] = verts[ex + a, ey + b, ez + c]
    # lattice nodes and their coordinates (from the trilinear map of a cell containing them)
    Ig, Jg, Kg = np.meshgrid(np.arange(Lx), np.arange(Ly), np.arange(Lz), indexing="ij")
    Ig = Ig.transpose(2, 1, 0).reshape(-1)
    Jg = Jg.transpose(2, 1, 0).reshape(-1)
    Kg = Kg.transpose(2, 1, 0).reshape(-1)
    cx = np.minimum(Ig // k, nx - 1)
    cy = np.minimum(Jg // k, ny - 1)
    cz = np.minimum(Kg // k, nz - 1)
    xi = np.stack([(Ig - k * cx) / k, (Jg - k * cy) / k, (Kg - k * cz) / k], axis=1)
    node_xyz = np.zeros((nnodes, 3))
    for v in range(8):
        a, b, c = v & 1, (v >> 1) & 1, (v >> 2) & 1
        w = (xi[:, 0] if a else 1 - xi[:, 0]) * (xi[:, 1] if b else 1 - xi[:, 1]) * (xi[:, 2] if c else 1 - xi[:, 2])
        node_xyz += w[:, None] * verts[cx + a, cy + b, cz + c]
    isdir = np.asarray(dirichlet(Ig, Jg, Kg, L), dtype=bool)
    free_rank = np.cumsum(~isdir) - 1
    dir_rank = np.cumsum(isdir) - 1
    nfree_nodes = int((~isdir).sum())
    ndir_nodes = int(isdir.sum())
    # signed ids per (node, comp): free 3*rank + c + 1, Dirichlet -(3*rank + c + 1)
    base = np.where(isdir, -(3 * dir_rank + 1), 3 * free_rank + 1).astype(np.int64)
    # local tensor nodes of each cell
    order_t = table_node_order(Q)  # order_t[n] = t
    nodeperm = np.empty(Q ** 3, dtype=np.int32)
    nodeperm[order_t] = np.arange(Q ** 3, dtype=np.int32)
    nn = Q ** 3
    ldof = np.array([[c * nn + n for n in range(nn)] for c in range(3)], dtype=np.int32)
    ti = order_t % Q
    tj = (order_t // Q) % Q
    tk = order_t // (Q * Q)
    gnode = ((k * ex)[:, None] + ti[None, :]) + Lx * (((k * ey)[:, None] + tj[None, :]) + Ly * ((k * ez)[:, None] + tk[None, :]))
    b = base[gnode]  # (ncells, nn)
    sgn = np.where(b > 0, 1, -1)
    cell_dofs = np.empty((ncells, 3 * nn), dtype=np.int64)
    for c in range(3):
        cell_dofs[:, ldof[c]] = b + sgn * c
    return HexSpace(n=(nx, ny, nz), order=k, domain=tuple(domain), cell_dofs=cell_dofs.astype(np.int32),
                    ldof=ldof, nodeperm=nodeperm, coords=coords, nfree=3 * nfree_nodes, ndir=3 * ndir_nodes,
                    node_xyz=node_xyz, node_free=~isdir, free_node_of=np.nonzero(~isdir)[0],
                    dir_node_of=np.nonzero(isdir)[0], lattice=L)


def boundary_faces(space: HexSpace, axis: int, side: int):
    """The cell faces on the box face ``axis`` = min (side 0) or max (side 1), the boundary entity a Gridap tag
    names (e.g. the x = 1 face): (face_cell, face_local) with the cells numbered x fastest and the local face
    2 axis + side of the reference cube."""
    nx, ny, nz = space.n
    I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    cell = (I + nx * (J + ny * K)).transpose(2, 1, 0).reshape(-1)
    idx = (I, J, K)[axis].transpose(2, 1, 0).reshape(-1)
    on = idx == (0 if side == 0 else space.n[axis] - 1)
    face_cell = cell[on].astype(np.int32)
    return face_cell, np.full(face_cell.size, 2 * axis + side, dtype=np.int32)


def interior_facets(space: HexSpace):
    """The interior facets of the Cartesian mesh (a SkeletonTriangulation's, the cells numbered x fastest): for
    every pair of cells sharing a face, (plus cell, plus local face, minus cell, minus local face, orientation) as
    int32 arrays -- the plus cell below along the facet's normal axis with its face 2 axis + 1, the minus cell above
    with its face 2 axis; both cells have the same local axes, so every orientation is 0."""
    nx, ny, nz = space.n
    cell = lambda i, j, k: i + nx * (j + ny * k)
    out = [[], [], [], [], []]
    for axis in range(3):
        for k in range(nz):
            for j in range(ny):
                for i in range(nx):
                    ijk = [i, j, k]
                    if ijk[axis] + 1 >= space.n[axis]:
                        continue
                    nb = list(ijk)
                    nb[axis] += 1
                    for lst, v in zip(out, (cell(*ijk), 2 * axis + 1, cell(*nb), 2 * axis, 0)):
                        lst.append(v)
    return tuple(np.asarray(a, dtype=np.int32) for a in out)


def _trilinear(X, xi):
    w = [(xi[0] if v & 1 else 1 - xi[0]) * (xi[1] if v & 2 else 1 - xi[1]) * (xi[2] if v & 4 else 1 - xi[2])
         for v in range(8)]
    return np.einsum("v,vi->i", np.asarray(w), np.asarray(X))


def _face_xi(lf, s, t):
    d, side = lf >> 1, lf & 1
    e1, e2 = (1, 2) if d == 0 else ((0, 2) if d == 1 else (0, 1))
    xi = [0.0, 0.0, 0.0]
    xi[d], xi[e1], xi[e2] = float(side), s, t
    return xi


def facet_orientation(coords_plus, face_plus, coords_minus, face_minus) -> int:
    """The orientation code (0..7) under which the minus face's coordinates follow the plus face's:
    (s', t') = (t, s) if o & 1 else (s, t), then s' -> 1 - s' if o & 2, t' -> 1 - t' if o & 4 (the corners of the
    two faces of the cells' trilinear maps are matched); raises when the faces do not coincide."""
    for o in range(8):
        ok = True
        for s, t in ((0, 0), (1, 0), (0, 1), (1, 1)):
            s2, t2 = (t, s) if o & 1 else (s, t)
            s2, t2 = (1 - s2 if o & 2 else s2), (1 - t2 if o & 4 else t2)
            xp = _trilinear(coords_plus, _face_xi(face_plus, s, t))
            xm = _trilinear(coords_minus, _face_xi(face_minus, s2, t2))
            if not np.allclose(xp, xm, atol=1e-12):
                ok = False
                break
        if ok:
            return o
    raise ValueError("HyperFEMKernels: the two faces do not coincide")


def refinement(fine: HexSpace, coarse: HexSpace):
    """(parent, child) of every cell of ``fine``, whose Cartesian mesh halves every cell of ``coarse`` (n_fine =
    2 n_coarse on the same box): fine cell (i, j, k) lies in coarse cell (i//2, j//2, k//2) as child
    (i%2) + 2 (j%2) + 4 (k%2) -- the arrays hfk_op_set_refinement takes (Gridap's refinement glue on a Julia host)."""
    nf = tuple(int(v) for v in fine.n)
    nc = tuple(int(v) for v in coarse.n)
    if any(a != 2 * b for a, b in zip(nf, nc)) or tuple(fine.domain) != tuple(coarse.domain):
        raise ValueError("HyperFEMKernels: the fine mesh must halve every cell of the coarse one on the same box")
    i, j, k = np.meshgrid(np.arange(nf[0]), np.arange(nf[1]), np.arange(nf[2]), indexing="ij")
    i, j, k = (a.transpose(2, 1, 0).reshape(-1) for a in (i, j, k))  # cells lexicographic, x fastest
    parent = (i // 2) + nc[0] * ((j // 2) + nc[1] * (k // 2))
    child = (i % 2) + 2 * (j % 2) + 4 * (k % 2)
    return parent.astype(np.int32), child.astype(np.int32)


def h_hierarchy(space: HexSpace, levels: int, dirichlet=None):
    """Q1 spaces on the meshes of ``space`` coarsened ``levels`` times by 2 in every direction (finest first, the
    first on space's own mesh), with the same Dirichlet rule: the h-levels below a p-hierarchy's Q1 level."""
    out = [hex_space(space.n, 1, domain=space.domain, dirichlet=dirichlet)]
    for _ in range(levels):
        n = out[-1].n
        if any(int(v) % 2 for v in n):
            raise ValueError(f"HyperFEMKernels: the mesh {tuple(n)} cannot be halved")
        out.append(hex_space(tuple(int(v) // 2 for v in n), 1, domain=space.domain, dirichlet=dirichlet))
    return out
