"""Pins of the oracle's operator-level functions on tiny meshes (brute force and invariants).

* residual = gradient of the total energy (central differences), with Dirichlet values;
* point gradients of affine fields = their constant gradient on trilinear cells;
* boundary traction load: total force t x area and first moment t x area x centroid per component;
* Robin face mass: symmetric, k x area per component, k int x^2 dGamma for u = x e_c, no component coupling;
* element Jacobians assembled densely = derivative of the residual (central differences);
* sum-factorised stored-T product (src/matrixfree.jl:341-424) = dense assembled matrix times dx
  (test/test_matrixfree.jl:74-96), serial = threaded (coloured determinism, test/test_assembly.jl:40-47);
* pattern (src/assembly.jl:773-899) = Python set of (row, col) pairs; positions consistent;
* assembled values through nzpos = dense matrix entries;
* colouring (src/coloring.jl:4-28): no two cells of a colour share a free dof, and each cell has the smallest
  colour not used by an earlier cell it shares a free dof with (greedy first fit);
* geometry: sum of dV = volume; affine patch test (linear law, affine field: interior residual 0);
* DOF counts of the benchmark meshes (open Dirichlet faces x = 0 and x = 1).
"""
import numpy as np
import pytest

from hfk.mesh import dirichlet_closed_x0, dirichlet_none, hex_space


def dense_matrix(space, Ke):
    n = space.nfree
    K = np.zeros((n, n))
    for cell in range(space.ncells):
        ids = space.cell_dofs[cell]
        f = ids > 0
        r = ids[f] - 1
        K[np.ix_(r, r)] += Ke[cell][np.ix_(f, f)]
    return K


def setup(order, n=(2, 2, 2), law="svk", perturb=0.0, dirichlet=None, seed=0):
    kw = {} if dirichlet is None else {"dirichlet": dirichlet}
    sp = hex_space(n, order, perturb=perturb, seed=seed, **kw)
    rng = np.random.default_rng(seed + 10)
    dv = 0.02 * rng.standard_normal(sp.ndir)
    x = 0.02 * rng.standard_normal(sp.nfree)
    return sp, dv, x


@pytest.mark.parametrize("order,law", [(1, "svk"), (2, "neohookean"), (2, "svk"), (3, "linear")])
def test_residual_is_energy_gradient(oracle, order, law):
    n = (2, 1, 1) if order == 3 else (2, 2, 1)
    sp, dv, x = setup(order, n, law, perturb=0.15)
    op = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    b = op.residual(x)
    h = 1e-6
    rng = np.random.default_rng(0)
    idx = rng.choice(sp.nfree, size=min(40, sp.nfree), replace=False)
    for i in idx:
        e = np.zeros(sp.nfree)
        e[i] = h
        fd = (op.energy(x + e) - op.energy(x - e)) / (2 * h)
        assert fd == pytest.approx(b[i], abs=1e-9, rel=1e-6)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_point_gradients_of_affine_fields(oracle, order):
    """An affine displacement u = A x + b is held exactly by Q1-Q3 on trilinear cells: its gradient at every
    quadrature point is A (row c, column J = d u_c / d x_J), with the field's Dirichlet values."""
    sp = hex_space((2, 2, 1), order, perturb=0.15, seed=3)
    A = np.array([[0.1, -0.2, 0.05], [0.3, 0.0, -0.1], [0.02, 0.15, -0.25]])

    def u(X):
        return X @ A.T + np.array([0.5, -1.0, 0.25])

    op = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=sp.dirichlet_values(u))
    G = op.point_gradients(sp.free_values(u))
    assert G.shape == (sp.ncells, (order + 1) ** 3, 3, 3)
    assert np.abs(G - A).max() < 1e-13


@pytest.mark.parametrize("order", [1, 2, 3])
def test_face_load_closed_forms(oracle, order):
    """A dead traction on two box faces with different tractions: per component the load sums to t x area (the
    face's functions sum to one) and its first moment sum_n f_n x_n to t x area x centroid (the nodes reproduce the
    trilinear map); interior vertices perturbed (boundary faces stay planar), Dirichlet dofs included."""
    from hfk.mesh import boundary_faces

    sp = hex_space((3, 2, 2), order, domain=(0.0, 2.0, 0.0, 1.5, 0.0, 0.5), perturb=0.15, seed=1,
                   dirichlet=dirichlet_closed_x0)
    op = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=np.zeros(sp.ndir))
    fc1, fl1 = boundary_faces(sp, 0, 1)
    fc2, fl2 = boundary_faces(sp, 1, 0)
    t1, t2 = np.array([0.3, -0.2, 0.7]), np.array([-1.1, 0.4, 0.25])
    tr = np.concatenate([np.tile(t1, (fc1.size, 1)), np.tile(t2, (fc2.size, 1))])
    ff, fd = op.face_load(np.concatenate([fc1, fc2]), np.concatenate([fl1, fl2]), tr)
    A1, A2 = 1.5 * 0.5, 2.0 * 0.5
    m1, m2 = A1 * np.array([2.0, 0.75, 0.25]), A2 * np.array([1.0, 0.0, 0.25])
    f = np.concatenate([ff, fd])
    comp = np.concatenate([np.arange(sp.nfree) % 3, np.arange(sp.ndir) % 3])
    xyz = np.concatenate([sp.free_xyz(), np.repeat(sp.node_xyz[sp.dir_node_of], 3, axis=0)])
    for c in range(3):
        sel = comp == c
        assert f[sel].sum() == pytest.approx(t1[c] * A1 + t2[c] * A2, rel=1e-13)
        assert np.allclose((f[sel][:, None] * xyz[sel]).sum(axis=0), t1[c] * m1 + t2[c] * m2, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_face_mass_closed_forms(oracle, order):
    """The Robin term on two box faces with stiffnesses k1, k2: symmetric; each component block sums to
    k1 A1 + k2 A2 (the face functions sum to one); for u = x e_c (held exactly) u.M u = k1 int x^2 + k2 int x^2 over
    the faces (x = Lx: Lx^2 A1; z = 0: Lx^3/3 Ly); no coupling between components."""
    from hfk.mesh import boundary_faces

    sp = hex_space((3, 2, 2), order, domain=(0.0, 2.0, 0.0, 1.5, 0.0, 0.5), perturb=0.15, seed=1,
                   dirichlet=dirichlet_closed_x0)
    op = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=np.zeros(sp.ndir))
    fc1, fl1 = boundary_faces(sp, 0, 1)
    fc2, fl2 = boundary_faces(sp, 2, 0)
    k1, k2 = 1.3, 0.4
    M = op.face_mass(np.concatenate([fc1, fc2]), np.concatenate([fl1, fl2]),
                     np.concatenate([np.full(fc1.size, k1), np.full(fc2.size, k2)]))
    assert np.abs(M - M.T).max() < 1e-15
    comp = np.concatenate([np.arange(sp.nfree) % 3, np.arange(sp.ndir) % 3])
    xyz = np.concatenate([sp.free_xyz(), np.repeat(sp.node_xyz[sp.dir_node_of], 3, axis=0)])
    for c in range(3):
        sel = comp == c
        assert M[np.ix_(sel, sel)].sum() == pytest.approx(k1 * 1.5 * 0.5 + k2 * 2.0 * 1.5, rel=1e-13)
        assert np.abs(M[np.ix_(sel, ~sel)]).max() == 0.0
        u = np.where(sel, xyz[:, 0], 0.0)
        assert u @ M @ u == pytest.approx(k1 * 4.0 * 0.75 + k2 * 8.0 / 3.0 * 1.5, rel=1e-13)


@pytest.mark.parametrize("order,law", [(1, "neohookean"), (2, "svk"), (2, "linear")])
def test_jacobian_is_residual_derivative(oracle, order, law):
    sp, dv, x = setup(order, (2, 2, 1), law, perturb=0.1)
    op = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    K = dense_matrix(sp, op.element_jacobians(x))
    h = 1e-6
    rng = np.random.default_rng(1)
    for j in rng.choice(sp.nfree, size=15, replace=False):
        e = np.zeros(sp.nfree)
        e[j] = h
        fd = (op.residual(x + e) - op.residual(x - e)) / (2 * h)
        assert np.allclose(K[:, j], fd, atol=1e-8, rtol=1e-6)


@pytest.mark.parametrize("order,P,law", [(1, 2, "svk"), (2, 3, "neohookean"), (2, 4, "svk"), (3, 4, "neohookean")])
def test_sum_factorised_equals_assembled(oracle, order, P, law):
    n = (2, 2, 1) if order == 3 else (3, 2, 2)
    sp, dv, x = setup(order, n, law, perturb=0.12)
    op = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    K = dense_matrix(sp, op.element_jacobians(x))
    prep = op.prepare(x)
    dx = np.random.default_rng(2).standard_normal(sp.nfree)
    y1 = op.jv(prep, dx, nthreads=1)
    y4 = op.jv(prep, dx, nthreads=4)
    assert np.array_equal(y1, y4)  # colour order: bitwise identical for any thread count
    ref = K @ dx
    assert np.linalg.norm(y1 - ref) / np.linalg.norm(ref) < 1e-12


def test_pattern_brute_force(oracle):
    sp, dv, x = setup(2, (3, 2, 2))
    op = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=dv)
    colptr, rowval, nzpos = op.pattern()
    pairs = set()
    for cell in range(sp.ncells):
        ids = sp.cell_dofs[cell]
        f = ids[ids > 0]
        for c in f:
            for r in f:
                pairs.add((int(c), int(r)))
    assert rowval.size == len(pairs)
    got = set()
    for c in range(sp.nfree):
        rows = rowval[colptr[c] - 1:colptr[c + 1] - 1]
        assert np.all(np.diff(rows) > 0)  # sorted, unique
        got.update((c + 1, int(r)) for r in rows)
    assert got == pairs
    nd = sp.nd
    for cell in range(sp.ncells):
        ids = sp.cell_dofs[cell]
        for jj in range(nd):
            for ii in range(nd):
                p = nzpos[cell, ii + nd * jj]
                if ids[ii] > 0 and ids[jj] > 0:
                    c = ids[jj]
                    assert colptr[c - 1] <= p < colptr[c] and rowval[p - 1] == ids[ii]
                else:
                    assert p == 0
    # assembled values = dense matrix entries
    K = dense_matrix(sp, op.element_jacobians(x))
    nz = op.jacobian(x, nzpos, rowval.size)
    for c in range(sp.nfree):
        for p in range(colptr[c] - 1, colptr[c + 1] - 1):
            assert nz[p] == pytest.approx(K[rowval[p] - 1, c], abs=1e-13)


@pytest.mark.parametrize("order", [1, 2])
def test_colouring_greedy_first_fit(oracle, order):
    sp, dv, x = setup(order, (4, 3, 3))
    col = oracle.OracleOp(sp, "svk", 2.0, 1.0).colors()
    cells_of = {}
    for cell in range(sp.ncells):
        for d in sp.cell_dofs[cell]:
            if d > 0:
                cells_of.setdefault(int(d), []).append(cell)
    for cell in range(sp.ncells):
        earlier = {col[o] for d in sp.cell_dofs[cell] if d > 0 for o in cells_of[int(d)] if o < cell}
        expect = 1
        while expect in earlier:
            expect += 1
        assert col[cell] == expect
    for d, cells in cells_of.items():
        assert len({col[c] for c in cells}) == len(cells)
    assert col.max() == 8  # a structured hex mesh needs 8 colours


def test_volume_and_patch(oracle):
    # sum dV over all points is the volume; a linear field under the linear law has zero interior residual
    sp = hex_space((3, 2, 2), 2, perturb=0.15, seed=3, dirichlet=lambda I, J, K, L: (I == 0) | (I == L[0] - 1) |
                   (J == 0) | (J == L[1] - 1) | (K == 0) | (K == L[2] - 1))
    A = np.array([[0.01, 0.02, -0.01], [0.0, 0.015, 0.005], [0.01, -0.02, 0.0]])
    f = lambda xyz: xyz @ A.T
    op = oracle.OracleOp(sp, "linear", 2.0, 1.0, dvals=sp.dirichlet_values(f))
    T, iJt, dV = op.prepare(sp.free_values(f))
    assert dV.sum() == pytest.approx(1.0, abs=1e-13)
    b = op.residual(sp.free_values(f))
    assert np.abs(b).max() < 1e-14


def test_benchmark_dof_counts():
    sp = hex_space((20, 10, 10), 2)
    assert (sp.nfree, sp.nfree + sp.ndir) == (52077, 54243)
    sp = hex_space((8, 8, 8), 1, dirichlet=dirichlet_closed_x0)
    assert sp.ndir == 3 * 81
    sp = hex_space((2, 2, 2), 3, dirichlet=dirichlet_none)
    assert sp.nfree == 3 * 7 ** 3
