"""TEST INFRASTRUCTURE ONLY -- the oracle's restatement of nodal block Jacobi on an assembled matrix, used by tests/
to check libhfk's hfk_sys_* path and the host's block numbering (hfk.assembled).  Never imported by the product.

* ``field_blocks``: the reference's numbering (src/preconditioners.jl:151-171): per field, over cells in order and
  table nodes in order, the tuple of the field's component ids at the node is a key; a key with every component
  fixed makes no block; blocks are numbered in first-seen order; fixed components are slot 0.
* ``coupled_blocks``: the same walk with one key over every field's slots at the node; a key whose primary-field
  slots are all fixed but which has a free dof elsewhere (a wall vertex: velocity prescribed, pressure free) is
  set aside and, in first-seen order, its free dofs go to the first block (walking the cells that contain that key
  in cell order, their table nodes in order) that has a free primary dof and enough empty slots; none: a block of
  its own.
* ``extract``, ``invert``, ``apply``: block_diagonal! on an assembled matrix, invert_blocks! (empty slots are
  identity rows and columns; a non-finite block or inverse is counted) and _apply_blocks!
  (src/preconditioners.jl:284-340).
* ``sweeps``: damped block Jacobi z <- z + omega Binv (r - A z) from z = 0.

Pure Python loops over cells (small meshes only); the matrix is a scipy CSC matrix.
"""
from __future__ import annotations

import numpy as np

BLOCK = 6


def _keys(cell_dofs, node_table, slots, cell, node):
    out = []
    for s in slots:
        l = int(node_table[node][s])
        d = int(cell_dofs[cell][l]) if l >= 0 else 0
        out.append(d if d > 0 else 0)
    return tuple(out)


def field_blocks(cell_dofs, node_table, slot_field):
    """(nblocks, 6) int32: field 0's blocks, then field 1's, ... (src/preconditioners.jl:151-171)."""
    out = []
    for f in sorted(set(int(v) for v in slot_field)):
        slots = [s for s in range(len(slot_field)) if slot_field[s] == f]
        seen = {}
        for cell in range(len(cell_dofs)):
            for node in range(len(node_table)):
                key = _keys(cell_dofs, node_table, slots, cell, node)
                if all(d <= 0 for d in key):
                    continue  # every component fixed: no block
                if key not in seen:
                    seen[key] = len(seen)
        for key in sorted(seen, key=seen.get):
            out.append(list(key) + [0] * (BLOCK - len(key)))
    return np.array(out, dtype=np.int32).reshape(-1, BLOCK)


def coupled_blocks(cell_dofs, node_table, slot_field):
    """(nblocks, 6) int32: one block per node over every field, wall-node orphans attached (module docstring)."""
    slots = list(range(len(slot_field)))
    prim = [s for s in slots if slot_field[s] == min(slot_field)]
    ncell, nn = len(cell_dofs), len(node_table)
    seen, blocks, block_at = {}, [], {}
    orphans, orphan_cells = {}, {}
    for cell in range(ncell):
        for node in range(nn):
            key = _keys(cell_dofs, node_table, slots, cell, node)
            if all(d <= 0 for d in key):
                continue
            if any(key[s] > 0 for s in prim):
                if key not in seen:
                    seen[key] = len(blocks)
                    blocks.append(list(key) + [0] * (BLOCK - len(key)))
                block_at[(cell, node)] = seen[key]
            else:
                if key not in orphans:
                    orphans[key] = len(orphans)
                orphan_cells.setdefault(key, []).append(cell)
    extra = []
    for key in sorted(orphans, key=orphans.get):
        dofs = [d for d in key if d > 0]
        target = None
        for cell in orphan_cells[key]:
            for node in range(nn):
                b = block_at.get((cell, node))
                if b is not None and sum(1 for d in blocks[b] if d == 0) >= len(dofs):
                    target = b
                    break
            if target is not None:
                break
        if target is None:
            extra.append(dofs + [0] * (BLOCK - len(dofs)))
            continue
        for d in dofs:
            blocks[target][blocks[target].index(0)] = d
    return np.array(blocks + extra, dtype=np.int32).reshape(-1, BLOCK)


def extract(A, blocks):
    """B[b][i][j] = A[d_i, d_j] (0 outside the pattern), identity on empty slots."""
    A = A.tocsc()
    nb = len(blocks)
    B = np.zeros((nb, BLOCK, BLOCK))
    for b in range(nb):
        for i in range(BLOCK):
            for j in range(BLOCK):
                di, dj = int(blocks[b][i]), int(blocks[b][j])
                B[b, i, j] = A[di - 1, dj - 1] if di > 0 and dj > 0 else (1.0 if i == j else 0.0)
    return B


def invert(B):
    """Binv per block and the number of blocks that are not finite or whose inverse is not finite."""
    Binv = np.array(B, copy=True)
    bad = 0
    for b in range(B.shape[0]):
        if not np.all(np.isfinite(B[b])):
            bad += 1
            continue
        try:
            Bi = np.linalg.inv(B[b])
        except np.linalg.LinAlgError:
            Bi = np.full((BLOCK, BLOCK), np.nan)
        Binv[b] = Bi
        if not np.all(np.isfinite(Bi)):
            bad += 1
    return Binv, bad


def apply(Binv, blocks, r):
    """z[d_i] = sum_j Binv[i][j] r[d_j] over the slots of each block (src/preconditioners.jl:326-340)."""
    z = np.zeros_like(np.asarray(r, dtype=np.float64))
    for b in range(len(blocks)):
        for i in range(BLOCK):
            di = int(blocks[b][i])
            if di <= 0:
                continue
            s = 0.0
            for j in range(BLOCK):
                dj = int(blocks[b][j])
                if dj > 0:
                    s += Binv[b, i, j] * r[dj - 1]
            z[di - 1] = s
    return z


def sweeps(A, Binv, blocks, r, k, omega):
    """k damped block-Jacobi steps z <- z + omega Binv (r - A z) from z = 0."""
    z = omega * apply(Binv, blocks, r)
    for _ in range(1, k):
        z = z + omega * apply(Binv, blocks, r - A @ z)
    return z


def block_diagonal(blocks, B, n):
    """The block-diagonal matrix of the blocks (empty slots dropped), scipy CSC."""
    import scipy.sparse as ssp

    rows, cols, vals = [], [], []
    for b in range(len(blocks)):
        for i in range(BLOCK):
            for j in range(BLOCK):
                di, dj = int(blocks[b][i]), int(blocks[b][j])
                if di > 0 and dj > 0:
                    rows.append(di - 1)
                    cols.append(dj - 1)
                    vals.append(B[b, i, j])
    return ssp.csc_matrix((vals, (rows, cols)), shape=(n, n))


# ---- Vanka patches and restricted additive Schwarz ----------------------------------------------------------

STAR = ((0, 4, 5, 6), (1, 4, 7, 8), (2, 5, 7, 9), (3, 6, 8, 9))  # P2 nodes of vertex f and of its edges in a cell


def vanka_patches(cell_dofs):
    """Patches of a P2-P1 layout [c*10 + node | 30 + vertex]: per free pressure (first seen over (cell, vertex)) the
    pressure, then the free velocity dofs of its star (the vertex and the midpoints of its edges) in first-seen order
    over (cell, vertex, star node, component); when the star has no free velocity dof (a vertex where two walls
    meet), the free velocity dofs of every node of the cells holding the vertex, in first-seen order over (cell,
    node, component).  A velocity dof is owned by the first patch listing it.  Returns a list of (dofs, owned)."""
    order, members, cells_of = [], {}, {}
    for cell in range(len(cell_dofs)):
        for f in range(4):
            p = int(cell_dofs[cell][30 + f])
            if p <= 0:
                continue
            if p not in members:
                order.append(p)
                members[p] = []
            cells_of.setdefault(p, []).append(cell)
            for node in STAR[f]:
                for c in range(3):
                    d = int(cell_dofs[cell][c * 10 + node])
                    if d > 0 and d not in members[p]:
                        members[p].append(d)
    for p in order:
        if members[p]:
            continue
        for cell in cells_of[p]:
            for node in range(10):
                for c in range(3):
                    d = int(cell_dofs[cell][c * 10 + node])
                    if d > 0 and d not in members[p]:
                        members[p].append(d)
    owner = {}
    patches = []
    for k, p in enumerate(order):
        dofs = [p] + members[p]
        own = [1] + [0] * len(members[p])
        for i, d in enumerate(members[p]):
            if d not in owner:
                owner[d] = k
                own[i + 1] = 1
        patches.append((dofs, own))
    return patches


def ras_apply(A, patches, r):
    """z[d] = (A_p^-1 r_p)[d] from the patch p that owns d (dense inverses by numpy)."""
    A = A.tocsc()
    z = np.zeros(A.shape[0])
    for dofs, own in patches:
        ix = np.array(dofs) - 1
        Ap = A[ix][:, ix].toarray()
        y = np.linalg.solve(Ap, r[ix])
        for i, d in enumerate(dofs):
            if own[i]:
                z[d - 1] = y[i]
    return z
