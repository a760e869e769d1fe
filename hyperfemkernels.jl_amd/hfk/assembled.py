"""Assembled multi-field systems on the device (``hfk_sys_*``): a square CSC matrix (the values of ``jacobian!``,
e.g. the Oseen P2-P1 Jacobian) with nodal block Jacobi, damped block-Jacobi sweeps and GMRES / FGMRES.

The reference's BlockJacobi numbers one set of nodal blocks per field in first-seen order over (cell, table node)
(``field_blocks``, src/preconditioners.jl:151-171): the components of that field at a node.  A mixed formulation
whose pressure has no diagonal (Taylor-Hood without PSPG) leaves those pressure blocks singular, which is why the
reference's p-multigrid stays single-field (src/pmultigrid.jl:252-254).  ``node_blocks`` gives both groupings:

* ``"fields"``: the reference's blocks, field by field;
* ``"coupled"``: one block per node over every field's free dofs at it (velocity and pressure at a vertex), in the
  same first-seen order; a node whose primary-field dofs are all fixed (a wall vertex: velocity prescribed, pressure
  free) hands its free dofs to the first block, in (cell, table node) order over the cells that contain it, that
  has a free primary-field dof and room left (an edge node of the same cell), so that no block is pressure alone.

Blocks hold at most 6 dofs (a node's velocity and up to three attached pressures); empty slots are 0 and act as
the identity.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, ptr
from .solvers import KRYLOV, SolveResult

BLOCK = 6  # dofs per block (padded)


def node_blocks(cell_dofs, node_table, slot_field, mode: str = "coupled") -> np.ndarray:
    """Nodal blocks (nblocks, 6) int32 of free dofs (1-based, 0 = empty slot).

    cell_dofs (ncells, nd) signed ids (> 0 free); node_table (nnodes, nslots) the cell-local dof of each slot of
    each table node (-1: the node has no such slot); slot_field (nslots,) the field of each slot (field 0 is the
    primary one).  Blocks are numbered first-seen over (cell, table node), as field_blocks does."""
    cd = np.asarray(cell_dofs)
    tab = np.asarray(node_table)
    fld = np.asarray(slot_field)
    if tab.shape[1] > BLOCK:
        raise ValueError("HyperFEMKernels: node_blocks: at most 6 dofs per node")
    ncells, nn = cd.shape[0], tab.shape[0]
    keys = np.where(tab[None, :, :] >= 0, cd[:, np.maximum(tab, 0)], 0)  # (ncells, nn, nslots)
    keys = np.where(keys > 0, keys, 0).astype(np.int64)
    if mode == "fields":
        out = []
        for f in np.unique(fld):
            cols = np.nonzero(fld == f)[0]
            k = keys[:, :, cols].reshape(-1, cols.size)
            out.append(_first_seen(k[(k > 0).any(axis=1)]))
        blocks = np.zeros((sum(b.shape[0] for b in out), BLOCK), dtype=np.int32)
        row = 0
        for b in out:
            blocks[row:row + b.shape[0], :b.shape[1]] = b
            row += b.shape[0]
        return blocks
    if mode != "coupled":
        raise ValueError("HyperFEMKernels: node_blocks: mode must be 'coupled' or 'fields'")
    prim = fld == fld.min()
    flat = keys.reshape(-1, keys.shape[2])
    has_prim = (flat[:, prim] > 0).any(axis=1)
    has_any = (flat > 0).any(axis=1)
    pflat = flat[has_prim]
    first, inv = _row_groups(pflat)
    rank = np.empty(first.size, dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(first.size)
    blocks_k = pflat[first[np.argsort(first, kind="stable")]]
    blocks = np.zeros((blocks_k.shape[0], BLOCK), dtype=np.int32)
    blocks[:, :blocks_k.shape[1]] = blocks_k
    orphan = has_any & ~has_prim  # only non-primary dofs free at this node
    if not orphan.any():
        return _packed(blocks)
    # block of each (cell, node) with a primary dof, and the first (cell, node) of each orphan key
    block_of = np.full(flat.shape[0], -1, dtype=np.int64)
    order = rank[inv]
    block_of[np.nonzero(has_prim)[0]] = order
    okeys, ofirst = _unique_first(flat[orphan])
    oidx = np.nonzero(orphan)[0]
    cells_of = {}
    for row in oidx:  # cells containing each orphan key, in cell order
        cells_of.setdefault(tuple(flat[row]), []).append(row // nn)
    free_slots = (blocks == 0).sum(axis=1)
    extra = []
    for j in np.argsort(ofirst, kind="stable"):
        key = tuple(okeys[j])
        dofs = [d for d in key if d > 0]
        target = -1
        for cell in cells_of[key]:
            for node in range(nn):
                b = block_of[cell * nn + node]
                if b >= 0 and free_slots[b] >= len(dofs):
                    target = b
                    break
            if target >= 0:
                break
        if target < 0:
            extra.append(dofs)
            continue
        for d in dofs:
            slot = int(np.nonzero(blocks[target] == 0)[0][0])
            blocks[target, slot] = d
            free_slots[target] -= 1
    if extra:
        add = np.zeros((len(extra), BLOCK), dtype=np.int32)
        for i, d in enumerate(extra):
            add[i, :len(d)] = d
        blocks = np.concatenate([blocks, add])
    return _packed(blocks)


def _packed(blocks):
    return np.ascontiguousarray(blocks, dtype=np.int32)


def _row_groups(k):
    """Distinct rows of k: (first, inv) with first[g] the index of the first occurrence of distinct row g and
    inv[i] the group of row i.  A stable lexicographic sort of the rows (np.unique over whole rows sorts them as
    byte strings, seconds for the 5M rows at 40^3 x 6 tetrahedra); equal rows keep their order within a group, so the
    first of every group is its first occurrence."""
    k = np.asarray(k, dtype=np.int64)
    if k.shape[0] == 0:
        return np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64)
    bits = max(int(k.max()).bit_length(), 1)
    if bits * k.shape[1] <= 63:  # the whole row in one non-negative int64 key
        pk = np.zeros(k.shape[0], dtype=np.int64)
        for c in range(k.shape[1]):
            pk = (pk << bits) | k[:, c]
        o = np.argsort(pk, kind="stable")
        ks = pk[o]
        new = np.ones(ks.shape[0], dtype=bool)
        new[1:] = ks[1:] != ks[:-1]
    else:  # two 31-bit ids per key (dof ids are int32)
        if k.shape[1] % 2:
            k = np.concatenate([k, np.zeros((k.shape[0], 1), dtype=np.int64)], axis=1)
        pk = (k[:, 0::2] << 31) | k[:, 1::2]
        o = np.lexsort(pk.T[::-1])
        ks = pk[o]
        new = np.ones(ks.shape[0], dtype=bool)
        new[1:] = (ks[1:] != ks[:-1]).any(axis=1)
    inv = np.empty(k.shape[0], dtype=np.int64)
    inv[o] = np.cumsum(new) - 1
    return o[new], inv


def _unique_first(k):
    """Distinct rows of k and the index of their first occurrence."""
    first, _ = _row_groups(k)
    return np.asarray(k)[first], first


def _first_seen(k):
    """Distinct rows of k in the order of their first occurrence."""
    u, first = _unique_first(k)
    return u[np.argsort(first, kind="stable")]


def _first_seen_index(k):
    """For every row of k, the first-seen rank of its distinct value."""
    first, inv = _row_groups(k)
    rank = np.empty(first.size, dtype=np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(first.size)
    return rank[inv]


def taylor_hood_table():
    """The node table of the P2-P1 cell layout of hfk_oseen_create (library order): node f = 0..9 carries the
    velocity dofs [c*10 + f], the vertices f = 0..3 also the pressure [30 + f]; slot fields (u, u, u, p)."""
    tab = np.full((10, 4), -1, dtype=np.int64)
    for f in range(10):
        tab[f, :3] = [f, 10 + f, 20 + f]
        if f < 4:
            tab[f, 3] = 30 + f
    return tab, np.array([0, 0, 0, 1])


def taylor_hood_blocks(space, mode: str = "coupled") -> np.ndarray:
    """Nodal blocks of a Taylor-Hood P2-P1 space (``hfk.tetmesh``; library local order)."""
    tab, fld = taylor_hood_table()
    cd = space.cell_dofs
    perm = getattr(space, "local_perm", None)
    if perm is not None:  # a host order: back to the library's (cell_dofs[:, perm[i]] is library dof i)
        cd = cd[:, perm]
    return node_blocks(cd, tab, fld, mode)


STAR = ((0, 4, 5, 6), (1, 4, 7, 8), (2, 5, 7, 9), (3, 6, 8, 9))  # P2 nodes of vertex f and of its 3 edges in a cell


def taylor_hood_vanka_patches(space):
    """Vanka patches of a Taylor-Hood P2-P1 space (library local order), for restricted additive Schwarz.

    One patch per free pressure vertex, in first-seen order over (cell, vertex): the pressure dof, then the free
    velocity dofs of its star (the vertex and the midpoints of its edges) in first-seen order over (cell, vertex,
    star node, component) among the cells that hold the vertex.  Where two walls meet, a vertex's star can hold no
    free velocity dof while the pressure still couples to the velocity of its cells: such a patch takes the free
    velocity dofs of every node of those cells instead (first seen over (cell, node, component)).  A velocity dof is
    owned by the first patch that lists it, a pressure by its own patch; free dofs that no patch lists get a patch per
    node of their own.  Returns (ptr (npatch + 1,) int64, dofs int32, owned int8)."""
    cd = np.asarray(space.cell_dofs)
    perm = getattr(space, "local_perm", None)
    if perm is not None:
        cd = cd[:, perm]
    ncells = cd.shape[0]
    pdof = cd[:, 30:34].reshape(-1)  # (cell, vertex) in order
    star = np.array(STAR)
    vel = np.stack([cd[:, c * 10 + star] for c in range(3)], axis=-1).reshape(ncells * 4, 12)
    keep_p = pdof > 0
    pr, pfirst = np.unique(pdof[keep_p], return_index=True)
    order = np.argsort(pfirst, kind="stable")
    prank = np.empty(pr.size, dtype=np.int64)
    prank[order] = np.arange(pr.size)
    rows = np.nonzero(keep_p)[0]
    pid_rows = prank[np.searchsorted(pr, pdof[rows])]
    pid = np.repeat(pid_rows, 12)
    okey = (rows[:, None] * 12 + np.arange(12)[None, :]).reshape(-1)
    vd = vel[rows].reshape(-1)
    m = vd > 0
    pid, okey, vd = pid[m], okey[m], vd[m]
    # first occurrence of each (patch, dof) pair, then patch-major in that order
    pair = pid * (int(cd.max()) + 1) + vd
    _, first = np.unique(pair, return_index=True)
    srt = first[np.lexsort((okey[first], pid[first]))]
    pid, vd = pid[srt], vd[srt]
    npatch = pr.size
    bounds = np.searchsorted(pid, np.arange(npatch + 1))
    members = [vd[bounds[k]:bounds[k + 1]] for k in range(npatch)]
    empty = [k for k in range(npatch) if members[k].size == 0]
    if empty:  # the velocity of every node of the cells holding the vertex
        emp = np.asarray(empty, dtype=np.int64)
        sel = np.isin(pid_rows, emp)
        kk, cc = pid_rows[sel], (rows // 4)[sel]
        srt = np.argsort(kk, kind="stable")  # the cells of each patch stay in cell order
        kk, cc = kk[srt], cc[srt]
        lo, hi = np.searchsorted(kk, emp, side="left"), np.searchsorted(kk, emp, side="right")
        vel_of = np.stack([cd[:, c * 10:(c + 1) * 10] for c in range(3)], axis=-1)  # (cell, node, c)
        for i, k in enumerate(empty):
            v = vel_of[cc[lo[i]:hi[i]]].reshape(-1)
            v = v[v > 0]
            _, f = np.unique(v, return_index=True)
            members[k] = v[np.sort(f)]
    lists = [np.concatenate([[pr[order[k]]], members[k]]) for k in range(npatch)]
    n = int(space.nfree)
    listed = np.zeros(n + 1, dtype=bool)
    for l in lists:
        listed[l] = True
    rest = np.nonzero(~listed[1:])[0] + 1
    if rest.size:  # free dofs outside every star: nodal patches (velocity triples, or single dofs)
        tab, fld = taylor_hood_table()
        nb = node_blocks(cd, tab, fld, "fields")
        lists += [b[b > 0] for b in nb if (b > 0).any() and np.isin(b[b > 0], rest).all()]
    ptr = np.zeros(len(lists) + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([l.size for l in lists])
    dofs = np.concatenate(lists).astype(np.int64)
    owned = np.zeros(dofs.size, dtype=np.int8)
    owned[np.unique(dofs, return_index=True)[1]] = 1  # patch-major: the first listing is the first patch
    return ptr, np.ascontiguousarray(dofs, dtype=np.int32), owned


class AssembledSystem:
    """A square assembled matrix on the device with nodal block Jacobi, damped block-Jacobi sweeps and GMRES
    (``hfk_sys_*``).  ``colptr``/``rowval`` the 1-based CSC pattern (e.g. ``OseenOperator.pattern()``)."""

    def __init__(self, colptr, rowval, ctx=None):
        from .operator import default_context

        self.ctx = ctx or default_context()
        self.lib = self.ctx.lib
        colptr = np.ascontiguousarray(colptr, dtype=np.int64)
        rowval = np.ascontiguousarray(rowval, dtype=np.int64)
        self.n = colptr.size - 1
        self.nnz = rowval.size
        h = C.c_void_p()
        check(self.lib.hfk_sys_create(self.ctx.handle, self.n, ptr(colptr, C.c_int64), ptr(rowval, C.c_int64),
                                      C.byref(h)), self.ctx.handle)
        self.handle = h
        self.nblocks = 0

    def set_blocks(self, blocks):
        b = np.ascontiguousarray(blocks, dtype=np.int32)
        if b.ndim != 2 or b.shape[1] != BLOCK:
            raise ValueError("HyperFEMKernels: blocks must be (nblocks, 6)")
        check(self.lib.hfk_sys_set_blocks(self.handle, b.shape[0], ptr(b, C.c_int32)), self.ctx.handle)
        self.nblocks = b.shape[0]
        return self

    def update_dev(self, nz_ptr: int, stream: int = 0, strict: bool = False) -> int:
        """New values (device CSC values of nnz doubles): the product copy and the inverted blocks.  Returns the
        number of blocks whose inverse is not finite; ``strict`` raises instead (the reference's strict mode)."""
        bad = C.c_int64()
        check(self.lib.hfk_sys_update_dev(self.handle, C.c_void_p(nz_ptr), C.byref(bad), C.c_void_p(stream)),
              self.ctx.handle)
        if strict and bad.value:
            raise np.linalg.LinAlgError(f"HyperFEMKernels: BlockJacobi: {bad.value} singular nodal blocks")
        return bad.value

    # host-array conveniences (device copies through torch)
    def _dev(self, a):
        import torch

        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(f"cuda:{self.ctx.device}")

    def update(self, values, strict: bool = False) -> int:
        self._values = self._dev(values)
        self._sync()
        return self.update_dev(self._values.data_ptr(), strict=strict)

    def _sync(self):
        import torch

        torch.cuda.synchronize(self.ctx.device)

    def matvec(self, x):
        xd = self._dev(x)
        y = xd.new_empty(self.n)
        self._sync()
        check(self.lib.hfk_sys_matvec_dev(self.handle, C.c_void_p(xd.data_ptr()), C.c_void_p(y.data_ptr()), None),
              self.ctx.handle)
        self.ctx.synchronize()
        return y.cpu().numpy()

    def apply_blocks(self, r):
        rd = self._dev(r)
        z = rd.new_empty(self.n)
        self._sync()
        check(self.lib.hfk_sys_blocks_apply_dev(self.handle, C.c_void_p(rd.data_ptr()), C.c_void_p(z.data_ptr()),
                                                None), self.ctx.handle)
        self.ctx.synchronize()
        return z.cpu().numpy()

    def smooth(self, r, sweeps: int = 1, omega: float = 1.0):
        rd = self._dev(r)
        z = rd.new_empty(self.n)
        self._sync()
        check(self.lib.hfk_sys_smooth_dev(self.handle, sweeps, omega, C.c_void_p(rd.data_ptr()),
                                          C.c_void_p(z.data_ptr()), None), self.ctx.handle)
        self.ctx.synchronize()
        return z.cpu().numpy()

    def set_patches(self, ptr, dofs, owned):
        """Overlapping patches (``taylor_hood_vanka_patches``) for the restricted additive Schwarz preconditioner."""
        pp = np.ascontiguousarray(ptr, dtype=np.int64)
        dofs = np.ascontiguousarray(dofs, dtype=np.int32)
        owned = np.ascontiguousarray(owned, dtype=np.int8)
        if dofs.size != owned.size or pp[-1] != dofs.size:
            raise ValueError("HyperFEMKernels: patches: ptr, dofs and owned disagree")
        check(self.lib.hfk_sys_set_patches(self.handle, pp.size - 1, pp.ctypes.data_as(C.POINTER(C.c_int64)),
                                           dofs.ctypes.data_as(C.POINTER(C.c_int32)),
                                           owned.ctypes.data_as(C.POINTER(C.c_int8))), self.ctx.handle)
        self.patches = (pp, dofs, owned)
        return self

    def apply_patches(self, r):
        rd = self._dev(r)
        z = rd.new_empty(self.n)
        self._sync()
        check(self.lib.hfk_sys_patches_apply_dev(self.handle, C.c_void_p(rd.data_ptr()), C.c_void_p(z.data_ptr()),
                                                 None), self.ctx.handle)
        self.ctx.synchronize()
        return z.cpu().numpy()

    def patch_inverses(self):
        """The dense inverse of every patch (list of m x m arrays)."""
        p = self.patches[0]
        m = np.diff(p)
        flat = np.zeros(int((m * m).sum()))
        check(self.lib.hfk_sys_patch_inverses(self.handle, flat.ctypes.data_as(C.POINTER(C.c_double))),
              self.ctx.handle)
        out, o = [], 0
        for k in m:
            out.append(flat[o:o + k * k].reshape(k, k))
            o += k * k
        return out

    def blocks(self):
        """(B, Binv): the extracted blocks and their inverses, (nblocks, 6, 6) each."""
        B = np.zeros((self.nblocks, BLOCK, BLOCK))
        Bi = np.zeros((self.nblocks, BLOCK, BLOCK))
        check(self.lib.hfk_sys_blocks(self.handle, ptr(B, C.c_double), ptr(Bi, C.c_double)), self.ctx.handle)
        return B, Bi

    def solve_dev(self, b_ptr: int, x_ptr: int, krylov: str = "gmres", preconditioner: str | None = "blocks",
                  sweeps: int = 2, omega: float = 0.8, rtol: float = 1e-8, atol: float = 0.0, maxiter: int = 1000,
                  restart: int = 50, stream: int = 0) -> SolveResult:
        pc = {None: 0, "none": 0, "blocks": 1, "sweeps": 2, "patches": 3}[preconditioner]
        kind = KRYLOV[krylov]
        st = np.zeros(3, dtype=np.int32)
        rn = np.zeros(2)
        check(self.lib.hfk_sys_solve_dev(self.handle, kind, pc, sweeps, omega, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                         rtol, atol, maxiter, restart, ptr(st, C.c_int32), ptr(rn, C.c_double),
                                         C.c_void_p(stream)), self.ctx.handle)
        return SolveResult(bool(st[0]), int(st[1]), bool(st[2]), float(rn[0]), float(rn[1]))

    def solve(self, b, **kw):
        bd = self._dev(b)
        x = bd.new_zeros(self.n)
        self._sync()
        res = self.solve_dev(bd.data_ptr(), x.data_ptr(), **kw)
        self.ctx.synchronize()
        return x.cpu().numpy(), res

    def close(self):
        if getattr(self, "handle", None):
            self.lib.hfk_sys_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
