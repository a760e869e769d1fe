"""TEST INFRASTRUCTURE ONLY -- the oracle's restatement of the Oseen-BDF2 step on tetrahedra (Taylor-Hood P2-P1,
SUPG/PSPG/LSIC), used by tests/ to check libhfk's hfk_oseen_* path.  Never imported by the product.

It evaluates the weak form of test/common.jl:165-228 (volume terms, the Laplacian in the stabilisation load, the
outlet face terms) literally, with tensor algebra on every
(trial, test) pair of basis functions -- grad u with Gridap's convention grad u[J,c] = d u_c/d x_J, eps(u),
(grad u)' . w, div -- instead of the per-entry reduced formulas of the device kernel.  Assembly follows the
reference: residual! (src/assembly.jl:909-916), sparsity as the sorted unique keys (col << 32) | row of the free
pairs of every cell (src/assembly.jl:773-818, 835-893) with nzpos by search (:864-899), greedy colouring over the
free test dofs (src/coloring.jl:4-28).  Numpy, for small meshes.
"""
from __future__ import annotations

import numpy as np

EDGES = ((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))
GL = np.array([[-1.0, -1.0, -1.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])  # reference grads of lambda


def basis(qp):
    """Reference values (np, 14) and gradients (np, 14, 3): P2 nodes 0..9 then P1 vertices 0..3."""
    qp = np.asarray(qp)
    lam = np.stack([1.0 - qp.sum(axis=1), qp[:, 0], qp[:, 1], qp[:, 2]], axis=1)
    npt = qp.shape[0]
    N = np.zeros((npt, 14))
    G = np.zeros((npt, 14, 3))
    for a in range(4):
        N[:, a] = lam[:, a] * (2 * lam[:, a] - 1)
        G[:, a] = (4 * lam[:, a] - 1)[:, None] * GL[a][None, :]
    for e, (a, b) in enumerate(EDGES):
        N[:, 4 + e] = 4 * lam[:, a] * lam[:, b]
        G[:, 4 + e] = 4 * (lam[:, b][:, None] * GL[a][None, :] + lam[:, a][:, None] * GL[b][None, :])
    N[:, 10:] = lam
    G[:, 10:] = GL[None, :, :]
    return N, G


def element_terms(coords, qp, qw, fields, rho, mu, c0, lap=None, tau=None):
    """Per cell, the matrices of every term of a((u,p),(v,q)) (dict name -> (ncells, 34 test, 34 trial)) and the
    load vector (ncells, 34).  fields (ncells, np, 9): w, un, f; lap (ncells, np, 3): the Laplacian of un in the
    stabilisation load rk = rho f + mu lap (test/common.jl:216-218), or None.  tau: the stabilisation parameters
    (ts, tp, tl) as a function of (|un|, h, rho, mu, c0) (arrays over cells and points); None for the reference's
    (test/common.jl:201-205)."""
    coords = np.asarray(coords)
    nc = coords.shape[0]
    Nr, Gr = basis(qp)
    J = np.transpose(coords[:, 1:, :] - coords[:, :1, :], (0, 2, 1))  # J[c][i][d] = x_{d+1,i} - x_{0,i}
    det = np.linalg.det(J)
    Jit = np.transpose(np.linalg.inv(J), (0, 2, 1))
    Gx = np.einsum("cij,qfj->cqfi", Jit, Gr)  # physical gradients (nc, np, 14, 3)
    dV = np.abs(det)[:, None] * np.asarray(qw)[None, :]
    vol = np.abs(det) / 6.0
    h = (6 * np.sqrt(2) * vol) ** (1.0 / 3.0)
    w, un, f = fields[..., 0:3], fields[..., 3:6], fields[..., 6:9]
    unn = np.linalg.norm(un, axis=2)
    if tau is None:
        ts = ((2 * unn / h[:, None]) ** 2 + 9 * (4 * mu / (rho * h[:, None] ** 2)) ** 2) ** -0.5
        tp = ts / rho
        tl = rho * unn * h[:, None] * 0.5
    else:
        ts, tp, tl = tau(unn, np.broadcast_to(h[:, None], unn.shape), rho, mu, c0)
    npt = len(qw)
    # the 34 functions: (u (3), grad u (3x3: [J, c]), p, grad p) at every point
    U = np.zeros((nc, npt, 34, 3))
    GU = np.zeros((nc, npt, 34, 3, 3))
    Pv = np.zeros((nc, npt, 34))
    GP = np.zeros((nc, npt, 34, 3))
    for c in range(3):
        for a in range(10):
            k = c * 10 + a
            U[:, :, k, c] = Nr[None, :, a]
            GU[:, :, k, :, c] = Gx[:, :, a, :]
    for a in range(4):
        Pv[:, :, 30 + a] = Nr[None, :, 10 + a]
        GP[:, :, 30 + a, :] = Gx[:, :, 10 + a, :]
    eps = 0.5 * (GU + np.swapaxes(GU, -1, -2))
    div = np.einsum("cqfjj->cqf", GU)
    conv = np.einsum("zqfjk,zqj->zqfk", GU, w)  # (grad u)' . w: component k = sum_J grad u[J, k] w_J
    Ru = rho * c0 * U + rho * conv
    T = {}
    T["mass"] = np.einsum("cq,cqai,cqbi->cab", dV, rho * c0 * U, U)
    T["viscous"] = np.einsum("cq,cqaij,cqbij->cab", dV, 2 * mu * eps, eps)
    T["convection"] = np.einsum("cq,cqbi,cqai->cab", dV, rho * conv, U)
    T["pressure"] = -np.einsum("cq,cqb,cqa->cab", dV, Pv, div)
    T["continuity"] = -np.einsum("cq,cqa,cqb->cab", dV, Pv, div)
    T["supg"] = np.einsum("cq,cq,cqai,cqbi->cab", dV, ts, conv, Ru) + np.einsum("cq,cq,cqai,cqbi->cab", dV, ts, conv, GP)
    T["pspg"] = -np.einsum("cq,cq,cqai,cqbi->cab", dV, tp, GP, Ru) - np.einsum("cq,cq,cqai,cqbi->cab", dV, tp, GP, GP)
    T["lsic"] = np.einsum("cq,cq,cqa,cqb->cab", dV, tl, div, div)
    rk = rho * f if lap is None else rho * f + mu * np.asarray(lap)
    load = (np.einsum("cq,cqi,cqai->ca", dV, rho * f, U) + np.einsum("cq,cq,cqai,cqi->ca", dV, ts, conv, rk)
            - np.einsum("cq,cq,cqai,cqi->ca", dV, tp, GP, rk))
    return T, load


def element_system(coords, qp, qw, fields, rho, mu, c0, lap=None, tau=None):
    T, load = element_terms(coords, qp, qw, fields, rho, mu, c0, lap, tau)
    return sum(T.values()), load


def outlet_terms(xo, xb, normals, rho, Pc, ZA):
    """Per outlet face, the terms of test/common.jl:212-213 and 219 over the velocity dofs of its cell:
    Kf (nfaces, 34, 34) = -rho negpart(w.n) u.v on dGb + ZA (u.n)(v.n) on dGo, and the load Lf (nfaces, 34) =
    -Pc v.n on dGo.  xo [face][q] = (reference coordinates in the cell, dS); xb the same with w appended."""
    xo, xb, nrm = np.asarray(xo), np.asarray(xb), np.asarray(normals)
    nf = xo.shape[0]
    No = basis(xo[..., :3].reshape(-1, 3))[0][:, :10].reshape(nf, xo.shape[1], 10)
    Nb = basis(xb[..., :3].reshape(-1, 3))[0][:, :10].reshape(nf, xb.shape[1], 10)
    # velocity test/trial functions as vectors: U[f, q, (c, a), i] = N_a delta_ci
    eye = np.eye(3)
    Uo = np.einsum("fqa,ci->fqcai", No, eye).reshape(nf, xo.shape[1], 30, 3)
    Ub = np.einsum("fqa,ci->fqcai", Nb, eye).reshape(nf, xb.shape[1], 30, 3)
    wn = np.einsum("fqi,fi->fq", xb[..., 4:7], nrm)
    negpart = 0.5 * (wn - np.abs(wn))
    Kf = np.zeros((nf, 34, 34))
    Kf[:, :30, :30] = (np.einsum("fq,fqai,fqbi->fab", -rho * negpart * xb[..., 3], Ub, Ub) +
                       ZA * np.einsum("fq,fqai,fi,fqbj,fj->fab", xo[..., 3], Uo, nrm, Uo, nrm))
    Lf = np.zeros((nf, 34))
    Lf[:, :30] = -Pc * np.einsum("fq,fqai,fi->fa", xo[..., 3], Uo, nrm)
    return Kf, Lf


def add_faces(K, load, face_cell, Kf, Lf):
    """Element matrices and loads with the face terms added to their cells."""
    K, load = K.copy(), load.copy()
    np.add.at(K, np.asarray(face_cell), Kf)
    np.add.at(load, np.asarray(face_cell), Lf)
    return K, load


def element_values(cell_dofs, x, dvals):
    ids = np.asarray(cell_dofs)
    xe = np.where(ids > 0, np.asarray(x)[np.maximum(ids, 1) - 1], np.asarray(dvals)[np.maximum(-ids, 1) - 1])
    return xe


REF_VERTS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])


def step_fields(cell_dofs, coords, qp, un, un_dir, unm1, unm1_dir, aw, bw, c1, c2):
    """The point data of a step built from the FE functions of the two previous steps as the reference's forms
    build them (test/common.jl:183-196): un, unm1 evaluated at the points through the P2 basis, w = aw un + bw unm1
    (wadv), f = c1 un - c2 unm1 (hist); the Laplacian by the reference's own route (:191-195): grad(un) . e_k
    interpolated into the discontinuous P1 space (its values at the 4 vertices of each cell), then the divergence
    of that P1 field.  Returns fields (ncells, np, 9) = w, un, f and lap (ncells, np, 3)."""
    ids = np.asarray(cell_dofs)[:, :30]
    coords = np.asarray(coords)
    ue = element_values(ids, un, un_dir).reshape(-1, 3, 10)
    ve = element_values(ids, unm1, unm1_dir).reshape(-1, 3, 10)
    N = basis(qp)[0][:, :10]
    u = np.einsum("qa,cka->cqk", N, ue)
    v = np.einsum("qa,cka->cqk", N, ve)
    fields = np.concatenate([aw * u + bw * v, u, c1 * u - c2 * v], axis=2)
    Jinv = np.linalg.inv(np.transpose(coords[:, 1:, :] - coords[:, :1, :], (0, 2, 1)))  # [cell][d][i]
    Gv = basis(REF_VERTS)[1][:, :10]  # reference gradients of the P2 functions at the vertices (4, 10, 3)
    gu = np.einsum("cka,mad,cdi->cmki", ue, Gv, Jinv)  # d un_k / d x_i at vertex m
    dlam = np.einsum("md,cdi->cmi", GL, Jinv)  # physical gradients of the P1 functions
    lapc = np.einsum("cmki,cmi->ck", gu, dlam)  # div of sum_m gu[m, k, :] lambda_m
    lap = np.repeat(lapc[:, None, :], np.asarray(qp).shape[0], axis=1)
    return fields, lap


def face_velocity(cell_dofs, face_cell, xref, un, un_dir, unm1, unm1_dir, aw, bw):
    """w = aw un + bw unm1 at face points given by reference coordinates in their cell, xref (nfaces, nq, 3)."""
    xref = np.asarray(xref)
    nf, nq = xref.shape[:2]
    ids = np.asarray(cell_dofs)[np.asarray(face_cell), :30]
    ue = element_values(ids, un, un_dir).reshape(nf, 3, 10)
    ve = element_values(ids, unm1, unm1_dir).reshape(nf, 3, 10)
    N = basis(xref.reshape(-1, 3))[0][:, :10].reshape(nf, nq, 10)
    return np.einsum("fqa,fka->fqk", N, aw * ue + bw * ve)


def residual(space, x, dvals, K, load):
    """b[free test] = sum over cells of (K x_e - l) (Dirichlet values lifted)."""
    ids = space.cell_dofs
    re = np.einsum("cab,cb->ca", K, element_values(ids, x, dvals)) - load
    b = np.zeros(space.nfree)
    m = ids > 0
    np.add.at(b, ids[m] - 1, re[m])
    return b


def pattern(cell_dofs, nfree):
    """The reference's sparsity: sorted unique keys (col << 32) | row over the free (test, trial) pairs of every
    cell; 1-based colptr/rowval; nzpos [cell][ii + nd jj] (0: not stored)."""
    ids = np.asarray(cell_dofs, dtype=np.int64)
    nc, nd = ids.shape
    rows = np.repeat(ids[:, :, None], nd, axis=2)  # [cell][ii][jj] -> test id
    cols = np.repeat(ids[:, None, :], nd, axis=1)  # trial id
    m = (rows > 0) & (cols > 0)
    keys = np.unique((cols[m] << 32) | rows[m])
    col = keys >> 32
    row = keys & 0xFFFFFFFF
    colptr = np.ones(nfree + 1, dtype=np.int64)
    np.add.at(colptr, col, 1)  # col is 1-based: counts land at index col
    colptr = np.cumsum(colptr) - np.arange(nfree + 1)  # 1 + cumulative counts
    colptr[0] = 1
    rowval = row.astype(np.int64)
    nzpos = np.zeros((nc, nd * nd), dtype=np.int32)
    cc = np.transpose(cols, (0, 2, 1)).reshape(nc, -1)  # index ii + nd jj -> [cell][jj][ii] order
    rr = np.transpose(rows, (0, 2, 1)).reshape(nc, -1)
    mm = (cc > 0) & (rr > 0)
    k = np.searchsorted(keys, (cc[mm] << 32) | rr[mm])
    nzpos[mm] = (k + 1).astype(np.int32)
    return colptr, rowval, nzpos


def block_pattern(cell_dofs, nfree, field_of, cell_mask):
    """The reference's pattern with field blocks (block_keys!, cell_positions! with FULL = false,
    src/assembly.jl:825-899): keys (col << 32) | row of the free pairs (ii, jj) whose block
    cell_mask[cell, field_of[ii], field_of[jj]] is set, sorted and unique; nzpos of every free pair is the position
    of its key when stored (by any cell), 0 otherwise."""
    ids = np.asarray(cell_dofs, dtype=np.int64)
    nc, nd = ids.shape
    fo = np.asarray(field_of)
    cm = np.asarray(cell_mask, dtype=bool)
    rows = np.repeat(ids[:, :, None], nd, axis=2)  # [cell][ii][jj]
    cols = np.repeat(ids[:, None, :], nd, axis=1)
    blk = cm[:, fo[:, None], fo[None, :]]  # [cell][ii][jj]
    m = (rows > 0) & (cols > 0) & blk
    keys = np.unique((cols[m] << 32) | rows[m])
    col = keys >> 32
    colptr = np.ones(nfree + 1, dtype=np.int64)
    np.add.at(colptr, col, 1)
    colptr = np.cumsum(colptr) - np.arange(nfree + 1)
    colptr[0] = 1
    rowval = (keys & 0xFFFFFFFF).astype(np.int64)
    nzpos = np.zeros((nc, nd * nd), dtype=np.int32)
    cc = np.transpose(cols, (0, 2, 1)).reshape(nc, -1)
    rr = np.transpose(rows, (0, 2, 1)).reshape(nc, -1)
    mm = (cc > 0) & (rr > 0)
    q = (cc[mm] << 32) | rr[mm]
    k = np.searchsorted(keys, q)
    found = (k < keys.size) & (keys[np.minimum(k, keys.size - 1)] == q)
    vals = np.where(found, k + 1, 0).astype(np.int32)
    nzpos[mm] = vals
    return colptr, rowval, nzpos


def jacobian_values(K, nzpos, nnz):
    nz = np.zeros(nnz)
    Kf = np.transpose(K, (0, 2, 1)).reshape(K.shape[0], -1)  # [cell][jj][ii] -> index ii + nd jj
    m = nzpos > 0
    np.add.at(nz, nzpos[m] - 1, Kf[m])
    return nz


def greedy_colors(cell_dofs, nfree):
    """src/coloring.jl:4-28: colour = lowest colour not used by any free test dof of the cell (first fit)."""
    used = [0] * nfree
    out = np.zeros(len(cell_dofs), dtype=np.int32)
    for c, ids in enumerate(np.asarray(cell_dofs)):
        mask = 0
        for i in ids:
            if i > 0:
                mask |= used[i - 1]
        k = 0
        while (mask >> k) & 1:
            k += 1
        out[c] = k + 1
        for i in ids:
            if i > 0:
                used[i - 1] |= 1 << k
    return out
