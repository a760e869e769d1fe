"""CPU pins of the oracle's Oseen-BDF2 restatement (tetrahedral P2-P1, SUPG/PSPG/LSIC; test/common.jl:165-228).

Each term of the element matrix is checked against what the mathematics fixes, on cells of a real mesh:
* basis: Kronecker property at the P2 nodes, partition of unity, exact reproduction of quadratics;
* quadrature: exact for every monomial of total degree <= 5;
* mass: the textbook P2 tetrahedron mass matrix |K|/420 [6, 1, -4, -6, 32, 16, 8];
* viscous 2 mu eps:eps: rigid motions (translations, rotations) are in its kernel, and it is not the full
  gradient form (a rotation is not in the kernel of grad:grad);
* pressure / continuity: a linear velocity field of trace t gives -t |K|/4 per pressure vertex;
* convection with constant w on a linear field u = A x: rho (A w)_i int N_a;
* LSIC: a divergence-free linear field is in its kernel; PSPG pressure block: -tp |K| grad M_a . grad M_b;
* assembly: the sparsity equals the set of coupled pairs; the greedy colours never share a free dof within a
  colour; the residual is affine, r(x) = J x + r(0);
* outlet faces: the triangle rules are exact to their degree; the faces cover the outlet (area, normal); with
  w.n = -1 the backflow block of each face is the textbook P2 face mass matrix |T|/180 [6, -1, 0, -4, 32, 16]; the
  Robin block sums to ZA |outlet| and the pressure load to -Pc |outlet|;
* Laplacian: the P2 elementwise Laplacian of quadratic fields; its load rows in closed form (PSPG
  -mu tp |K| lap . grad l_a; SUPG zero at the vertices, ts mu lap_i |K| w . (grad l_a + grad l_b) at edge (a, b)).
* step point data from FE functions: for quadratic u^n, u^{n-1} the point values, w, f, the Laplacian (by the
  reference's discontinuous-P1 route) and the outlet-face w equal the polynomials' closed forms.
"""
import numpy as np
import pytest

from hfk.tetmesh import EDGES, divfree_field, tet_p2p1_space, tet_quadrature_deg5
from oracle import oseen as O


@pytest.fixture(scope="module")
def mesh():
    sp = tet_p2p1_space((2, 2, 1), domain=(0.0, 1.0, 0.0, 0.8, 0.0, 0.6))
    qp, qw = tet_quadrature_deg5()
    return sp, qp, qw


def nodes_ref():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    E = np.array([(V[a] + V[b]) / 2 for a, b in EDGES])
    return np.vstack([V, E])


def test_p2_basis():
    X = nodes_ref()
    N, G = O.basis(X)
    assert np.allclose(N[:, :10], np.eye(10), atol=1e-15)
    rng = np.random.default_rng(0)
    pts = rng.dirichlet(np.ones(4), size=20)[:, 1:]
    N, G = O.basis(pts)
    assert np.allclose(N[:, :10].sum(axis=1), 1.0) and np.allclose(G[:, :10].sum(axis=1), 0.0)
    assert np.allclose(N[:, 10:].sum(axis=1), 1.0)
    # quadratic reproduction: f = 1 + 2x - y + 3xz + z^2
    f = lambda P: 1 + 2 * P[:, 0] - P[:, 1] + 3 * P[:, 0] * P[:, 2] + P[:, 2] ** 2
    assert np.allclose(N[:, :10] @ f(X), f(pts), atol=1e-14)


def test_quadrature_degree5():
    from math import factorial

    qp, qw = tet_quadrature_deg5()
    assert np.all(qw > 0) and np.all(qp >= 0) and np.all(qp.sum(axis=1) <= 1)
    for a in range(6):
        for b in range(6 - a):
            for c in range(6 - a - b):
                exact = factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)
                assert (qw * qp[:, 0] ** a * qp[:, 1] ** b * qp[:, 2] ** c).sum() == pytest.approx(exact, rel=1e-13)


def terms(sp, qp, qw, w=None, un=None, f=None, rho=1.3, mu=0.07, c0=20.0):
    npt = len(qw)
    fields = np.zeros((sp.ncells, npt, 9))
    for k, v in enumerate((w, un, f)):
        if v is not None:
            fields[:, :, 3 * k:3 * k + 3] = v
    return O.element_terms(sp.coords, qp, qw, fields, rho, mu, c0)


def test_mass_matrix_closed_form(mesh):
    sp, qp, qw = mesh
    T, _ = terms(sp, qp, qw, rho=1.3, c0=20.0)
    M = np.zeros((10, 10))
    for a in range(10):
        for b in range(10):
            if a < 4 and b < 4:
                M[a, b] = 6 if a == b else 1
            elif a < 4 or b < 4:
                v, e = (a, b - 4) if a < 4 else (b, a - 4)
                M[a, b] = -4 if v in EDGES[e] else -6
            else:
                ea, eb = EDGES[a - 4], EDGES[b - 4]
                M[a, b] = 32 if a == b else (16 if set(ea) & set(eb) else 8)
    vol = sp.volumes()
    for c in range(3):
        blk = T["mass"][:, c * 10:(c + 1) * 10, c * 10:(c + 1) * 10]
        assert np.allclose(blk, 1.3 * 20.0 * vol[:, None, None] * M[None] / 420, rtol=1e-12, atol=1e-16)
    # no coupling between components or with pressure
    assert np.abs(T["mass"][:, :10, 10:]).max() < 1e-16


def nodal_field(sp, cell_fun):
    """Element dof vectors (ncells, 34) of a velocity field given at the P2 nodes of every cell."""
    out = np.zeros((sp.ncells, 34))
    X = nodes_ref()
    for c in range(sp.ncells):
        x0 = sp.coords[c, 0]
        J = (sp.coords[c, 1:] - x0).T
        P = x0[None, :] + X @ J.T
        u = cell_fun(P)  # (10, 3)
        for k in range(3):
            out[c, k * 10:(k + 1) * 10] = u[:, k]
    return out


def test_viscous_kernel_rigid_motions(mesh):
    sp, qp, qw = mesh
    T, _ = terms(sp, qp, qw, mu=0.07)
    Kv = T["viscous"]
    for fun in (lambda P: np.tile([1.0, -2.0, 0.5], (len(P), 1)),               # translation
                lambda P: np.cross(np.array([0.3, -0.7, 1.1])[None, :], P)):    # rotation
        xe = nodal_field(sp, fun)
        assert np.abs(np.einsum("cab,cb->ca", Kv, xe)).max() < 1e-13 * np.abs(Kv).max()
    # a pure shear has energy 2 mu |eps|^2 |K|
    A = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    xe = nodal_field(sp, lambda P: P @ A.T)
    e = 0.5 * (A + A.T)
    energy = np.einsum("ca,cab,cb->c", xe, Kv, xe)
    assert np.allclose(energy, 2 * 0.07 * (e * e).sum() * sp.volumes(), rtol=1e-12)


def test_divergence_terms_linear_field(mesh):
    sp, qp, qw = mesh
    T, _ = terms(sp, qp, qw)
    A = np.array([[0.4, -0.2, 0.1], [0.3, 0.7, -0.5], [0.2, 0.1, -0.3]])
    xe = nodal_field(sp, lambda P: P @ A.T)
    tr = np.trace(A)
    q_rows = np.einsum("cab,cb->ca", T["continuity"], xe)[:, 30:]
    assert np.allclose(q_rows, -tr * sp.volumes()[:, None] / 4, rtol=1e-12)
    # the pressure term is its transpose
    assert np.allclose(T["pressure"], np.transpose(T["continuity"], (0, 2, 1)), atol=1e-16)


def test_convection_constant_w(mesh):
    sp, qp, qw = mesh
    w = np.array([0.5, -1.0, 2.0])
    T, _ = terms(sp, qp, qw, w=w, rho=1.3)
    A = np.array([[0.4, -0.2, 0.1], [0.3, 0.7, -0.5], [0.2, 0.1, -0.3]])
    xe = nodal_field(sp, lambda P: P @ A.T)
    out = np.einsum("cab,cb->ca", T["convection"], xe)
    intN = np.array([-1 / 20] * 4 + [1 / 5] * 6)  # int N_a / |K| of the P2 functions
    Aw = A @ w
    for i in range(3):
        assert np.allclose(out[:, i * 10:(i + 1) * 10], 1.3 * Aw[i] * sp.volumes()[:, None] * intN[None], rtol=1e-12)


def test_lsic_and_pspg_blocks(mesh):
    sp, qp, qw = mesh
    un = np.array([0.3, 0.4, 1.2])
    T, _ = terms(sp, qp, qw, w=un, un=un, rho=1.3, mu=0.07)
    A = np.array([[0.4, -0.2, 0.1], [0.3, 0.7, -0.5], [0.2, 0.1, -1.1]])  # trace 0: divergence free
    xe = nodal_field(sp, lambda P: P @ A.T)
    assert np.abs(np.einsum("cab,cb->ca", T["lsic"], xe)).max() < 1e-12 * np.abs(T["lsic"]).max()
    vol = sp.volumes()
    h = (6 * np.sqrt(2) * vol) ** (1 / 3)
    ts = ((2 * np.linalg.norm(un) / h) ** 2 + 9 * (4 * 0.07 / (1.3 * h ** 2)) ** 2) ** -0.5
    tp = ts / 1.3
    G = np.zeros((sp.ncells, 4, 3))
    for c in range(sp.ncells):
        J = (sp.coords[c, 1:] - sp.coords[c, 0]).T
        G[c] = O.GL @ np.linalg.inv(J)
    ref = -tp[:, None, None] * vol[:, None, None] * np.einsum("cai,cbi->cab", G, G)
    assert np.allclose(T["pspg"][:, 30:, 30:], ref, rtol=1e-12)


def test_pattern_colors_and_affine_residual(mesh):
    sp, qp, qw = mesh
    colptr, rowval, nzpos = O.pattern(sp.cell_dofs, sp.nfree)
    pairs = set()
    for ids in sp.cell_dofs:
        f = [int(i) for i in ids if i > 0]
        pairs.update((c, r) for c in f for r in f)
    col = np.repeat(np.arange(1, sp.nfree + 1), np.diff(colptr))
    assert set(zip(col.tolist(), rowval.tolist())) == pairs
    assert np.all(np.diff(rowval)[np.diff(col) == 0] > 0)  # sorted within columns
    colors = O.greedy_colors(sp.cell_dofs, sp.nfree)
    for k in range(1, colors.max() + 1):
        ids = sp.cell_dofs[colors == k]
        f = ids[ids > 0]
        assert f.size == np.unique(f).size
    rng = np.random.default_rng(3)
    fields = rng.standard_normal((sp.ncells, len(qw), 9))
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.01, 150.0)
    nz = O.jacobian_values(K, nzpos, rowval.size)
    import scipy.sparse as ssp

    Am = ssp.csc_matrix((nz, rowval - 1, colptr - 1), shape=(sp.nfree, sp.nfree))
    d = rng.standard_normal(sp.ndir)
    x = rng.standard_normal(sp.nfree)
    r0 = O.residual(sp, np.zeros(sp.nfree), d, K, load)
    r = O.residual(sp, x, d, K, load)
    assert np.linalg.norm(r - r0 - Am @ x) < 1e-13 * np.linalg.norm(r)


def test_divfree_field():
    w = divfree_field(5)
    x = np.random.default_rng(1).random((50, 3))
    h = 1e-5
    div = sum((w(x + h * e)[:, i] - w(x - h * e)[:, i]) / (2 * h) for i, e in enumerate(np.eye(3)))
    assert np.abs(div).max() < 1e-6 * np.abs(w(x)).max()


# ---- outlet terms and the Laplacian of the stabilisation load (test/common.jl:191-219) ----------------------

@pytest.fixture(scope="module")
def channel():
    from hfk.tetmesh import outlet_faces

    sp = tet_p2p1_space((3, 2, 2), domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    return sp, outlet_faces(sp)


@pytest.mark.parametrize("degree", [4, 5])
def test_triangle_quadrature(degree):
    from math import factorial

    from hfk.tetmesh import tri_quadrature

    p, w = tri_quadrature(degree)
    assert np.all(w > 0) and np.all(p >= 0) and np.all(p.sum(axis=1) <= 1)
    for i in range(degree + 1):
        for j in range(degree + 1 - i):
            exact = factorial(i) * factorial(j) / factorial(i + j + 2)
            assert (w * p[:, 0] ** i * p[:, 1] ** j).sum() == pytest.approx(exact, rel=1e-13)


def test_outlet_faces_geometry(channel):
    sp, of = channel
    assert len(of.cell) == 2 * 2 * 2  # two tetrahedra of each of the 2 x 2 boundary boxes
    assert of.area == pytest.approx(1.0, rel=1e-14) and of.xo[..., 3].sum() == pytest.approx(1.0, rel=1e-14)
    assert np.allclose(of.normal, [1.0, 0.0, 0.0], atol=1e-15)
    X = sp.coords[of.cell]
    phys = X[:, None, 0, :] + np.einsum("fqd,fdi->fqi", of.xo[..., :3], X[:, 1:, :] - X[:, :1, :])
    assert np.abs(phys[..., 0] - 2.0).max() < 1e-14


def face_nodes(sp, of):
    """Cell-local P2 nodes on each outlet face: the 3 vertices and the 3 edges joining them."""
    out = []
    for f, c in enumerate(of.cell):
        on = np.abs(sp.coords[c, :, 0] - 2.0) < 1e-12
        verts = [k for k in range(4) if on[k]]
        edges = [4 + e for e, (a, b) in enumerate(EDGES) if on[a] and on[b]]
        out.append((verts, edges))
    return out


def test_outlet_backflow_is_the_p2_face_mass_matrix(channel):
    sp, of = channel
    nf, nqb = of.xb.shape[:2]
    inflow = np.concatenate([of.xb, np.tile([-1.0, 0.0, 0.0], (nf, nqb, 1))], axis=2)  # w.n = -1
    Kf, Lf = O.outlet_terms(of.xo, inflow, of.normal, 1.3, 0.0, 0.0)
    assert np.abs(Lf).max() == 0.0
    for f, (verts, edges) in enumerate(face_nodes(sp, of)):
        P = sp.coords[of.cell[f]][verts]
        area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
        nodes = verts + edges
        M = np.zeros((6, 6))
        for r, a in enumerate(nodes):
            for s_, b in enumerate(nodes):
                if a < 4 and b < 4:
                    M[r, s_] = 6 if a == b else -1
                elif a >= 4 and b >= 4:
                    M[r, s_] = 32 if a == b else 16
                else:
                    v, e = (a, b - 4) if a < 4 else (b, a - 4)
                    M[r, s_] = 0 if v in EDGES[e] else -4
        for c in range(3):
            idx = [c * 10 + a for a in nodes]
            assert np.allclose(Kf[f][np.ix_(idx, idx)], 1.3 * area * M / 180, rtol=1e-12, atol=1e-15)
        others = [k for k in range(30) if k % 10 not in nodes]
        assert np.abs(Kf[f][others]).max() < 1e-15
    outflow = np.concatenate([of.xb, np.tile([1.0, 0.0, 0.0], (nf, nqb, 1))], axis=2)
    Kf, _ = O.outlet_terms(of.xo, outflow, of.normal, 1.3, 0.0, 0.0)
    assert np.abs(Kf).max() == 0.0


def test_outlet_robin_and_pressure_totals(channel):
    sp, of = channel
    nf, nqb = of.xb.shape[:2]
    xb = np.concatenate([of.xb, np.zeros((nf, nqb, 3))], axis=2)
    Kf, Lf = O.outlet_terms(of.xo, xb, of.normal, 1.0, 0.7, 2.3)
    # sum_a N_a = 1 on a face: the x-x block sums to ZA x area, the load to -Pc x area; n has no y, z part
    assert Kf[:, :10, :10].sum() == pytest.approx(2.3 * of.area, rel=1e-13)
    assert np.abs(Kf[:, 10:30, :]).max() == 0.0 and np.abs(Kf[:, :, 10:30]).max() == 0.0
    assert Lf[:, :10].sum() == pytest.approx(-0.7 * of.area, rel=1e-13)
    assert np.abs(Lf[:, 10:]).max() == 0.0


def test_p2_laplacian_quadratics():
    from hfk.tetmesh import p2_laplacian

    sp = tet_p2p1_space((2, 3, 2), domain=(0.0, 1.0, 0.0, 0.8, 0.0, 0.6))
    for fun, lap in ((lambda X: np.stack([X[:, 0] ** 2, X[:, 1] * X[:, 2], X[:, 0] ** 2 + X[:, 1] ** 2], 1), [2, 0, 4]),
                     (lambda X: np.stack([X[:, 0] * X[:, 1], X[:, 2] ** 2 - X[:, 0], 3 * X[:, 1] ** 2], 1), [0, 2, 6])):
        x = np.zeros(sp.nfree)
        x[:sp.nfree_u] = fun(sp.node_xyz[sp.free_node_of]).reshape(-1)
        assert np.abs(p2_laplacian(sp, x, sp.dirichlet_values(fun)) - lap).max() < 1e-11


def test_laplacian_load_closed_forms(mesh):
    sp, qp, qw = mesh
    npt = len(qw)
    lap = np.array([0.4, -1.1, 0.7])
    w = np.array([0.5, -1.0, 2.0])
    un = np.array([0.3, 0.4, 1.2])
    rho, mu = 1.3, 0.07
    fields = np.zeros((sp.ncells, npt, 9))
    fields[:, :, 3:6] = un
    L = np.tile(lap, (sp.ncells, npt, 1))
    _, l0 = O.element_terms(sp.coords, qp, qw, fields, rho, mu, 20.0, lap=L)
    vol = sp.volumes()
    h = (6 * np.sqrt(2) * vol) ** (1 / 3)
    ts = ((2 * np.linalg.norm(un) / h) ** 2 + 9 * (4 * mu / (rho * h ** 2)) ** 2) ** -0.5
    G = np.zeros((sp.ncells, 4, 3))
    for c in range(sp.ncells):
        G[c] = O.GL @ np.linalg.inv((sp.coords[c, 1:] - sp.coords[c, 0]).T)
    # w = 0, f = 0: no SUPG; the PSPG rows -tp int grad M_a . (mu lap) = -mu tp |K| lap . grad l_a
    assert np.abs(l0[:, :30]).max() == 0.0
    assert np.allclose(l0[:, 30:], -mu * (ts / rho)[:, None] * vol[:, None] * (G @ lap), rtol=1e-12)
    fields[:, :, 0:3] = w
    _, l1 = O.element_terms(sp.coords, qp, qw, fields, rho, mu, 20.0, lap=L)
    # SUPG rows ts int (grad N_a . w) mu lap_i: zero at the vertices (int grad N_a = 0), ts mu lap_i |K| w.(grad l_a
    # + grad l_b) at the midpoint of edge (a, b)
    for i in range(3):
        rows = l1[:, i * 10:(i + 1) * 10]
        assert np.abs(rows[:, :4]).max() < 1e-13 * np.abs(rows).max()
        for e, (a, b) in enumerate(EDGES):
            ref = ts * mu * lap[i] * vol * ((G[:, a] + G[:, b]) @ w)
            assert np.allclose(rows[:, 4 + e], ref, rtol=1e-11, atol=1e-15)


def test_step_fields_closed_forms(channel):
    """The point data of a step from two quadratic velocity fields p (u^n) and q (u^{n-1}), which P2 holds exactly:
    the point values are the polynomials at the physical points, w = 2p - q and f = c1 p - c2 q (BDF2), the
    Laplacian (the reference's route through the discontinuous P1 gradient) is the polynomials' own, and the face
    velocity at the outlet points is 2p - q there."""
    sp, of = channel
    qp, qw = tet_quadrature_deg5()
    A = np.array([[[1.0, 0.3, -0.2], [0.3, -0.5, 0.1], [-0.2, 0.1, 0.7]],
                  [[0.2, 0.0, 0.4], [0.0, 1.1, -0.3], [0.4, -0.3, -0.6]],
                  [[-0.8, 0.5, 0.0], [0.5, 0.2, 0.2], [0.0, 0.2, 0.9]]])
    B = np.array([[0.5, -1.0, 0.2], [0.0, 0.3, 1.0], [1.5, 0.1, -0.4]])

    def p(X):
        return np.einsum("cij,mi,mj->mc", A, X, X) + X @ B.T + np.array([0.1, -0.2, 0.3])

    def q(X):
        return np.stack([X[:, 1] * X[:, 2], X[:, 0] ** 2 - X[:, 2], 2.0 - X[:, 1] ** 2], 1)

    rng = np.random.default_rng(3)
    un = rng.standard_normal(sp.nfree)  # the pressure entries are not read
    unm1 = rng.standard_normal(sp.nfree)
    un[:sp.nfree_u] = p(sp.node_xyz[sp.free_node_of]).reshape(-1)
    unm1[:sp.nfree_u] = q(sp.node_xyz[sp.free_node_of]).reshape(-1)
    c1, c2 = 200.0, 50.0
    fields, lap = O.step_fields(sp.cell_dofs, sp.coords, qp, un, sp.dirichlet_values(p), unm1, sp.dirichlet_values(q),
                                2.0, -1.0, c1, c2)
    X = sp.quadrature_points(qp).reshape(-1, 3)
    P, Q = p(X).reshape(sp.ncells, -1, 3), q(X).reshape(sp.ncells, -1, 3)
    assert np.abs(fields[..., 3:6] - P).max() < 1e-13
    assert np.abs(fields[..., 0:3] - (2 * P - Q)).max() < 1e-13
    assert np.abs(fields[..., 6:9] - (c1 * P - c2 * Q)).max() < 1e-10
    assert np.abs(lap - 2.0 * np.trace(A, axis1=1, axis2=2)).max() < 1e-11
    wb = O.face_velocity(sp.cell_dofs, of.cell, of.xb[..., :3], un, sp.dirichlet_values(p), unm1,
                         sp.dirichlet_values(q), 2.0, -1.0)
    Xb = of.xb_phys.reshape(-1, 3)
    assert np.abs(wb - (2 * p(Xb) - q(Xb)).reshape(wb.shape)).max() < 1e-13


def test_block_pattern_against_sets():
    """The oracle's field-block pattern (block_keys!/cell_positions!, src/assembly.jl:825-899) against a Python set of
    (row, col) pairs: no pressure-pressure block anywhere (Taylor-Hood Oseen without PSPG), then the pressure-pressure
    block on a random half of the cells only; every free pair's nzpos addresses its CSC entry when some cell's blocks
    store it and is 0 otherwise; with every block set it is the full pattern."""
    sp = tet_p2p1_space((3, 2, 2))
    ids = sp.cell_dofs
    fo = np.array([0] * 30 + [1] * 4)
    rng = np.random.default_rng(3)
    no_pp = np.ones((sp.ncells, 2, 2), dtype=bool)
    no_pp[:, 1, 1] = False
    part = no_pp.copy()
    part[rng.random(sp.ncells) < 0.5, 1, 1] = True
    full = O.pattern(ids, sp.nfree)
    assert all(np.array_equal(a, b) for a, b in zip(O.block_pattern(ids, sp.nfree, fo, np.ones((sp.ncells, 2, 2))), full))
    for cm in (no_pp, part):
        colptr, rowval, nzpos = O.block_pattern(ids, sp.nfree, fo, cm)
        pairs = set()
        for c in range(sp.ncells):
            for ii in range(34):
                for jj in range(34):
                    if ids[c, ii] > 0 and ids[c, jj] > 0 and cm[c, fo[ii], fo[jj]]:
                        pairs.add((int(ids[c, ii]), int(ids[c, jj])))
        got = {(int(rowval[k]), j + 1) for j in range(sp.nfree) for k in range(colptr[j] - 1, colptr[j + 1] - 1)}
        assert got == pairs and rowval.size == len(pairs)
        assert rowval.size < full[1].size
        col_of = np.repeat(np.arange(sp.nfree), np.diff(colptr))
        for c in range(sp.ncells):
            for ii in range(34):
                for jj in range(34):
                    p = nzpos[c, ii + 34 * jj]
                    r, q = ids[c, ii], ids[c, jj]
                    if r > 0 and q > 0 and (int(r), int(q)) in pairs:
                        assert p > 0 and rowval[p - 1] == r and col_of[p - 1] + 1 == q
                    else:
                        assert p == 0
