"""The oracle's nodal blocks on an assembled matrix (oracle/blocks.py), pinned on the CPU, and the host's block
numbering (hfk.assembled) against it, bit for bit.

Pins of the oracle:
* every free dof is in exactly one block slot; a per-field velocity block holds the 3 components of one lattice node
  (tetmesh's free ids 3 rank + c + 1 of that node's rank), a pressure block one vertex; the counts are the free
  velocity nodes and the pressure vertices;
* a coupled block is the velocity of one node, with the pressure of the same point (a vertex) or, for a wall vertex
  whose velocity is prescribed, of a vertex of a cell that holds the node (distance <= the cell diagonal);
* the first block is the first node of cell 0 with a free dof (first-seen numbering);
* extraction from a matrix whose entries are a closed form of (row, column) gives that form, identity on empty
  slots; B Binv = I; a zero pressure block is counted;
* the block application equals a sparse solve with the block-diagonal matrix; one undamped sweep on a
  block-diagonal matrix is its exact inverse, and many damped sweeps on a block-dominant matrix reach A^-1 r.
"""
import numpy as np
import pytest
import scipy.sparse as ssp
import scipy.sparse.linalg as sla

from hfk.assembled import node_blocks, taylor_hood_blocks, taylor_hood_table
from hfk.tetmesh import tet_p2p1_space
from oracle import blocks as OB


def space(n=(3, 2, 2)):
    return tet_p2p1_space(n, domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))


def covers(blocks, n):
    d = blocks[blocks > 0]
    return np.array_equal(np.sort(d), np.arange(1, n + 1))


def test_field_blocks_pinned():
    sp = space()
    tab, fld = taylor_hood_table()
    B = OB.field_blocks(sp.cell_dofs, tab, fld)
    assert covers(B, sp.nfree)
    nvel = sp.nfree_u // 3
    assert B.shape[0] == nvel + sp.vert_xyz.shape[0]
    vel, pres = B[:nvel], B[nvel:]
    assert np.all(vel[:, 3:] == 0) and np.all(pres[:, 1:] == 0)
    rank = (vel[:, 0] - 1) // 3
    assert np.array_equal(vel[:, :3], np.stack([3 * rank + 1, 3 * rank + 2, 3 * rank + 3], axis=1))
    # first seen: the first block is the first node of cell 0 with a free velocity
    first = next(f for f in range(10) if sp.cell_dofs[0, f] > 0)
    assert vel[0, 0] == sp.cell_dofs[0, first]
    assert pres[0, 0] == sp.cell_dofs[0, 30]


def test_coupled_blocks_pinned():
    sp = space()
    tab, fld = taylor_hood_table()
    B = OB.coupled_blocks(sp.cell_dofs, tab, fld)
    assert covers(B, sp.nfree)
    nu = sp.nfree_u
    h = np.sqrt((2.0 / 3) ** 2 + 0.5 ** 2 + 0.5 ** 2)  # the cell diagonal of the (3, 2, 2) box mesh
    full = {int(d) for blk in B if np.count_nonzero(blk) == OB.BLOCK for d in blk}
    for blk in B:
        u = [d for d in blk if 0 < d <= nu]
        p = [d for d in blk if d > nu]
        assert len(u) in (0, 3) and len(p) <= (3 if u else 1)
        if not u:
            # a pressure alone: every cell holding that vertex has each free velocity node's block full
            cells = np.nonzero((sp.cell_dofs[:, 30:34] == p[0]).any(axis=1))[0]
            assert all(int(d) in full for c in cells for d in sp.cell_dofs[c, :30] if d > 0)
            continue
        node = sp.free_node_of[(u[0] - 1) // 3]
        assert u == [3 * ((u[0] - 1) // 3) + c + 1 for c in range(3)]
        for q in p:
            dist = np.linalg.norm(sp.node_xyz[node] - sp.vert_xyz[q - nu - 1])
            assert dist <= h + 1e-12
    # vertices with a free velocity keep their own pressure (distance 0)
    own = [blk for blk in B if blk[0] > 0 and blk[3] > 0 and np.allclose(
        sp.node_xyz[sp.free_node_of[(blk[0] - 1) // 3]], sp.vert_xyz[blk[3] - nu - 1])]
    interior = [v for v in range(sp.vert_xyz.shape[0])
                if not np.any(np.isclose(sp.vert_xyz[v, 1:], [0.0, 0.0]) | np.isclose(sp.vert_xyz[v, 1:], [1.0, 1.0]))
                and not np.isclose(sp.vert_xyz[v, 0], 0.0)]
    assert len(own) == len(interior)


@pytest.mark.parametrize("mode", ["fields", "coupled"])
def test_host_numbering_equals_oracle(mode):
    sp = space((3, 3, 2))
    tab, fld = taylor_hood_table()
    ref = (OB.field_blocks if mode == "fields" else OB.coupled_blocks)(sp.cell_dofs, tab, fld)
    assert np.array_equal(taylor_hood_blocks(sp, mode), ref)


def test_host_numbering_generic_tables():
    """node_blocks on a scalar + vector layout with partially fixed components, both modes, against the oracle."""
    rng = np.random.default_rng(3)
    ncell, nn = 7, 5
    tab = np.array([[0, 5, 10, 15], [1, 6, 11, -1], [2, 7, 12, 16], [3, 8, 13, -1], [4, 9, 14, 17]])
    fld = np.array([0, 0, 0, 1])
    ids = rng.permutation(200) + 1
    cd = np.zeros((ncell, 18), dtype=np.int64)
    pool = iter(ids)
    shared = {}
    for c in range(ncell):
        for s in range(18):
            key = (c // 2, s) if s % 3 else (c, s)  # some dofs shared between cells
            if key not in shared:
                shared[key] = next(pool)
            cd[c, s] = shared[key]
    cd[rng.random(cd.shape) < 0.25] *= -1  # fixed
    for mode, fn in (("fields", OB.field_blocks), ("coupled", OB.coupled_blocks)):
        assert np.array_equal(node_blocks(cd, tab, fld, mode), fn(cd, tab, fld))


def closed_form_matrix(n, density=0.05, seed=0):
    rng = np.random.default_rng(seed)
    M = ssp.random(n, n, density=density, random_state=rng, format="csc")
    M = (M + ssp.eye(n)).tocsc()
    M.sort_indices()
    r, c = M.nonzero()
    M = ssp.csc_matrix(((r + 1) + (c + 1) / 1000.0, (r, c)), shape=(n, n))
    return M


def test_extract_closed_form_and_inverse():
    n = 40
    A = closed_form_matrix(n)
    blocks = np.zeros((14, OB.BLOCK), dtype=np.int32)
    perm = np.random.default_rng(1).permutation(n) + 1
    k = 0
    for b in range(14):
        m = 2 if b < 2 else 3
        blocks[b, :m] = perm[k:k + m]
        k += m
    B = OB.extract(A, blocks)
    D = A.toarray()
    for b in range(14):
        for i in range(OB.BLOCK):
            for j in range(OB.BLOCK):
                di, dj = blocks[b, i], blocks[b, j]
                want = (D[di - 1, dj - 1] if di and dj else float(i == j))
                assert B[b, i, j] == want
                if di and dj and D[di - 1, dj - 1] != 0:
                    assert B[b, i, j] == di + dj / 1000.0
    Bi, bad = OB.invert(B)
    ok = [b for b in range(14) if abs(np.linalg.det(B[b])) > 1e-8]
    for b in ok:
        assert np.abs(B[b] @ Bi[b] - np.eye(OB.BLOCK)).max() < 1e-10
    assert bad == 14 - len(ok)
    Z = B.copy()
    Z[0, :2, :2] = 0.0  # a zero block (a pressure block without a diagonal)
    assert OB.invert(Z)[1] == bad + 1


def test_apply_and_sweeps():
    n = 30
    rng = np.random.default_rng(2)
    blocks = np.zeros((10, OB.BLOCK), dtype=np.int32)
    blocks[:, :3] = (rng.permutation(n) + 1).reshape(10, 3)
    Bd = np.zeros((10, OB.BLOCK, OB.BLOCK))
    for b in range(10):
        Bd[b] = np.eye(OB.BLOCK)
        Bd[b, :3, :3] = rng.standard_normal((3, 3)) + 4 * np.eye(3)
    D = OB.block_diagonal(blocks, Bd, n)
    Bi, bad = OB.invert(Bd)
    assert bad == 0
    r = rng.standard_normal(n)
    z = OB.apply(Bi, blocks, r)
    assert np.linalg.norm(z - sla.spsolve(D, r)) < 1e-12 * np.linalg.norm(z)
    assert np.linalg.norm(OB.sweeps(D, Bi, blocks, r, 1, 1.0) - z) < 1e-14 * np.linalg.norm(z)
    E = ssp.random(n, n, density=0.1, random_state=np.random.default_rng(5), format="csc") * 0.3
    A = (D + E).tocsc()
    Bi2, _ = OB.invert(OB.extract(A, blocks))
    x = OB.sweeps(A, Bi2, blocks, r, 300, 0.8)
    assert np.linalg.norm(x - sla.spsolve(A, r)) < 1e-9 * np.linalg.norm(x)


def test_vanka_patches_pinned():
    """Vanka patches: an interior vertex of the 6-tetrahedra box split has 14 edges, so its patch is the pressure
    and 3 x 15 velocity dofs; the velocity dofs of a patch all sit at the vertex or at the midpoints of its edges,
    or, where that star holds no free velocity (two walls meeting), none of them does and all lie in the cells
    around the vertex; every free dof is owned once; the host's patches equal the oracle's, dof for dof."""
    from hfk.assembled import taylor_hood_vanka_patches

    sp = space((3, 3, 2))
    pat = OB.vanka_patches(sp.cell_dofs)
    nu = sp.nfree_u
    owned = sorted(d for dofs, own in pat for d, o in zip(dofs, own) if o)
    assert owned == list(range(1, sp.nfree + 1))
    sizes = []
    for dofs, own in pat:
        v = sp.vert_xyz[dofs[0] - nu - 1]
        assert dofs[0] > nu and own[0] == 1 and all(d <= nu for d in dofs[1:]) and len(dofs) > 1
        in_star = []
        for d in dofs[1:]:
            x = sp.node_xyz[sp.free_node_of[(d - 1) // 3]]
            e = 2 * (x - v) / np.array([2.0 / 3, 1.0 / 3, 0.5])  # twice the offset in box steps
            assert np.allclose(e, np.round(e)) and np.all(np.abs(np.round(e)) <= 2)  # a node of a cell around v
            e = np.round(e)
            in_star.append(bool(np.all(np.abs(e) <= 1) and (np.all(e >= 0) or np.all(e <= 0))))
        assert all(in_star) or not any(in_star)
        interior = 0 < v[1] < 1 and 0 < v[2] < 1 and 0 < v[0] < 2
        sizes.append((interior, len(dofs)))
    assert {s for i, s in sizes if i} == {46}
    ptr, dofs, own = taylor_hood_vanka_patches(sp)
    assert len(ptr) - 1 == len(pat)
    for k, (d, o) in enumerate(pat):
        assert list(dofs[ptr[k]:ptr[k + 1]]) == d and list(own[ptr[k]:ptr[k + 1]]) == o


def test_ras_exact_cases():
    """Restricted additive Schwarz: one patch holding every dof is the exact solve; patches that do not overlap on a
    block-diagonal matrix are block Jacobi, also exact."""
    n = 12
    rng = np.random.default_rng(7)
    A = ssp.csc_matrix(rng.standard_normal((n, n)) + 5 * np.eye(n))
    r = rng.standard_normal(n)
    one = [(list(range(1, n + 1)), [1] * n)]
    assert np.linalg.norm(OB.ras_apply(A, one, r) - np.linalg.solve(A.toarray(), r)) < 1e-13 * np.linalg.norm(r)
    D = np.zeros((n, n))
    parts = [list(range(1, 5)), list(range(5, 10)), list(range(10, 13))]
    for p in parts:
        ix = np.array(p) - 1
        D[np.ix_(ix, ix)] = rng.standard_normal((len(p), len(p))) + 4 * np.eye(len(p))
    pats = [(p, [1] * len(p)) for p in parts]
    z = OB.ras_apply(ssp.csc_matrix(D), pats, r)
    assert np.linalg.norm(z - np.linalg.solve(D, r)) < 1e-13 * np.linalg.norm(z)
