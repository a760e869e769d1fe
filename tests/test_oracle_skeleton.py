"""The oracle's skeleton (interior facet) terms, pinned on the CPU (the device tests compare against them).

* Closed form: the gradient-jump penalty of a field with a kink on the plane x = 1/2 (u_c = max(x - 1/2, 0), exact
  in Q1-Q3) is (gamma / 2) * 3 * h * area summed over that plane's facets: 1.5 gamma / sqrt(ny nz) on the unit cube.
* Fields with a continuous gradient (a global quadratic on affine cells, a global linear field on perturbed cells)
  have no jump: residual and energy vanish.
* The residual is the gradient of the energy, the Jacobian the derivative of the residual (central differences),
  and symmetric; the product equals the matrix times the increment.
* The orientation: the same mesh with one cell's local axes rotated (facet orientation 5) gives the same energy,
  residual and matrix (up to the renumbering of nothing: the dofs are the same).
* The cross-block pairs are those of a set built from the facets' cells.
"""
import dataclasses

import numpy as np
import pytest

from hfk.mesh import dirichlet_none, facet_orientation, hex_space, interior_facets
from oracle import skeleton as S


def kinked(X):
    return np.repeat(np.maximum(X[:, :1] - 0.5, 0.0), 3, axis=1)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_kink_closed_form(order):
    sp = hex_space((2, 3, 2), order, dirichlet=dirichlet_none)
    fac = interior_facets(sp)
    x = sp.free_values(kinked)
    _, e, _ = S.skeleton_terms(sp, fac, S.GradJump, [0.7], x, None)
    assert e == pytest.approx(1.5 * 0.7 / np.sqrt(3 * 2), rel=1e-12)


def test_continuous_gradient_has_no_jump():
    sp = hex_space((2, 2, 2), 2, dirichlet=dirichlet_none)
    quad = lambda X: np.stack([X[:, 0] ** 2 + X[:, 1] * X[:, 2], X[:, 1] ** 2 - X[:, 0], X[:, 2] * X[:, 0]], axis=1)
    r, e, _ = S.skeleton_terms(sp, interior_facets(sp), S.GradJump, [1.0], sp.free_values(quad), None)
    assert np.abs(r).max() < 1e-13 and abs(e) < 1e-26
    spp = hex_space((2, 2, 2), 2, dirichlet=dirichlet_none, perturb=0.2, seed=3)
    lin = lambda X: X @ np.array([[1.0, 0.2, -0.3], [0.5, -1.0, 0.1], [0.0, 0.4, 2.0]])
    r, e, _ = S.skeleton_terms(spp, interior_facets(spp), S.GradJump, [1.0], spp.free_values(lin), None)
    assert np.abs(r).max() < 1e-12 and abs(e) < 1e-24


@pytest.mark.parametrize("form,params", [(S.GradJump, [0.8]), (S.NonlinearFacet, [0.8, 2.0, 0.6, 1.5])])
def test_residual_gradient_and_matrix(form, params):
    sp = hex_space((2, 2, 1), 2, dirichlet=dirichlet_none, perturb=0.2, seed=5)
    fac = interior_facets(sp)
    rng = np.random.default_rng(2)
    x = 0.05 * rng.standard_normal(sp.nfree)
    dx = rng.standard_normal(sp.nfree)
    r, e, y = S.skeleton_terms(sp, fac, form, params, x, None, dx=dx)
    A = S.skeleton_matrix(sp, fac, form, params, x, None)
    h = 1e-5
    for k in rng.choice(sp.nfree, 12, replace=False):  # r = dE/dx (central differences, O(h^2) ~ 1e-10 relative)
        xp, xm = x.copy(), x.copy()
        xp[k] += h
        xm[k] -= h
        ep = S.skeleton_terms(sp, fac, form, params, xp, None)[1]
        em = S.skeleton_terms(sp, fac, form, params, xm, None)[1]
        assert (ep - em) / (2 * h) == pytest.approx(r[k], rel=1e-6, abs=1e-8 * np.abs(r).max())
    rp = S.skeleton_terms(sp, fac, form, params, x + h * dx, None)[0]
    rm = S.skeleton_terms(sp, fac, form, params, x - h * dx, None)[0]
    fd = (rp - rm) / (2 * h)
    assert np.linalg.norm(fd - A @ dx) < 1e-7 * np.linalg.norm(fd)  # A = dr/dx
    assert np.linalg.norm(y - A @ dx) < 1e-13 * np.linalg.norm(y)  # the product is the matrix
    assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()  # a second derivative


def rotated_pair(order):
    """Two cells along x, the second with its local axes rotated about x (local xi -> standard (xi0, 1 - xi2, xi1)):
    the same mesh and dofs, another local numbering of the second cell."""
    sp = hex_space((2, 1, 1), order, dirichlet=dirichlet_none, perturb=0.0)
    Q = sp.Q
    ids = sp.cell_dofs[1].copy()
    X = sp.coords[1].copy()
    new_ids = np.zeros_like(ids)
    new_X = np.zeros_like(X)
    for t in range(Q ** 3):
        i, j, k = t % Q, (t // Q) % Q, t // (Q * Q)
        ts = i + Q * ((Q - 1 - k) + Q * j)
        for c in range(3):
            new_ids[sp.ldof[c, sp.nodeperm[t]]] = ids[sp.ldof[c, sp.nodeperm[ts]]]
    for v in range(8):
        b0, b1, b2 = v & 1, (v >> 1) & 1, (v >> 2) & 1
        vs = b0 + 2 * (1 - b2) + 4 * b1
        new_X[v] = X[vs]
    cd = sp.cell_dofs.copy()
    cd[1] = new_ids
    co = sp.coords.copy()
    co[1] = new_X
    rot = dataclasses.replace(sp, cell_dofs=cd, coords=co)
    o = facet_orientation(co[0], 1, co[1], 0)
    return sp, rot, o


@pytest.mark.parametrize("order", [1, 2])
def test_orientation_invariance(order):
    sp, rot, o = rotated_pair(order)
    assert o == 5
    fac0 = interior_facets(sp)
    fac1 = (fac0[0], fac0[1], fac0[2], np.array([0], dtype=np.int32), np.array([o], dtype=np.int32))
    rng = np.random.default_rng(4)
    x = rng.standard_normal(sp.nfree)
    p = [0.8, 2.0, 0.6, 1.5]
    r0, e0, _ = S.skeleton_terms(sp, fac0, S.NonlinearFacet, p, x, None)
    r1, e1, _ = S.skeleton_terms(rot, fac1, S.NonlinearFacet, p, x, None)
    assert e1 == pytest.approx(e0, rel=1e-12)
    assert np.linalg.norm(r1 - r0) < 1e-12 * np.linalg.norm(r0)
    A0 = S.skeleton_matrix(sp, fac0, S.NonlinearFacet, p, x, None)
    A1 = S.skeleton_matrix(rot, fac1, S.NonlinearFacet, p, x, None)
    assert np.abs(A1 - A0).max() < 1e-12 * np.abs(A0).max()
    with pytest.raises(ValueError):  # a face that is not the shared one
        facet_orientation(rot.coords[0], 1, rot.coords[1], 1)


def test_pairs_are_the_facet_cells():
    sp = hex_space((3, 2, 1), 1)
    fac = interior_facets(sp)
    pairs = S.skeleton_pairs(sp, fac)
    for f in range(len(fac[0])):
        a = {int(i) for i in sp.cell_dofs[fac[0][f]] if i > 0}
        b = {int(i) for i in sp.cell_dofs[fac[2][f]] if i > 0}
        assert all((c, r) in pairs for c in a for r in b)
    cells_of = {}
    for cell in range(sp.ncells):
        for i in sp.cell_dofs[cell]:
            if i > 0:
                cells_of.setdefault(int(i), set()).add(cell)
    nbr = {c: set() for c in range(sp.ncells)}
    for f in range(len(fac[0])):
        nbr[int(fac[0][f])].add(int(fac[2][f]))
        nbr[int(fac[2][f])].add(int(fac[0][f]))
    for c, r in pairs:  # every pair comes from two cells across a facet
        assert any(cr in nbr[cc] for cc in cells_of[c] for cr in cells_of[r])


# ---- boundary groups with a captured integrand (one cell per face) -------------------------------------------

def x1_faces(sp):
    from hfk.mesh import boundary_faces

    return boundary_faces(sp, 0, 1)


@pytest.mark.parametrize("u,force", [(lambda X: np.stack([0.2 * X[:, 0], 0 * X[:, 0], 0 * X[:, 0]], 1), (1.0, 0, 0)),
                                     (lambda X: np.stack([0 * X[:, 0], 0.3 * X[:, 1], 0 * X[:, 0]], 1), (1.3, 0, 0)),
                                     (lambda X: np.stack([0 * X[:, 0], 0 * X[:, 0], -0.25 * X[:, 2]], 1), (0.75, 0, 0))])
def test_follower_pressure_closed_form(u, force):
    """On the face x = 1 of the unit cube under a homogeneous deformation the total follower force is
    p cof(F) e_x |face|: stretching along x leaves it p e_x, stretching y or z by (1 + b) scales it by (1 + b)."""
    sp = hex_space((2, 2, 2), 2, dirichlet=dirichlet_none, perturb=0.2, seed=1)  # boundary vertices stay put
    r, e, _ = S.face_terms(sp, x1_faces(sp), S.FollowerPressure, [2.0], sp.free_values(u), None)
    tot = np.array([r[c::3].sum() for c in range(3)])  # free ids interleave the components
    assert np.allclose(tot, 2.0 * np.array(force), atol=1e-12) and e == 0.0


@pytest.mark.parametrize("form,params", [(S.FollowerPressure, [2.0]), (S.NonlinearBoundary, [0.7, 3.0, 0.4, 1.3])])
def test_boundary_form_derivatives(form, params):
    from hfk.mesh import boundary_faces

    sp = hex_space((2, 2, 1), 2, dirichlet=dirichlet_none, perturb=0.2, seed=2)
    faces = tuple(np.concatenate(a) for a in zip(boundary_faces(sp, 0, 1), boundary_faces(sp, 1, 0)))
    rng = np.random.default_rng(6)
    x = 0.05 * rng.standard_normal(sp.nfree)
    dx = rng.standard_normal(sp.nfree)
    r, e, y = S.face_terms(sp, faces, form, params, x, None, dx=dx)
    A = S.face_matrix(sp, faces, form, params, x, None)
    h = 1e-5
    rp = S.face_terms(sp, faces, form, params, x + h * dx, None)[0]
    rm = S.face_terms(sp, faces, form, params, x - h * dx, None)[0]
    fd = (rp - rm) / (2 * h)
    assert np.linalg.norm(fd - A @ dx) < 1e-7 * np.linalg.norm(fd)
    assert np.linalg.norm(y - A @ dx) < 1e-13 * np.linalg.norm(y)
    if form is S.NonlinearBoundary:  # conservative: r = dE/dx, a symmetric Jacobian
        for k in rng.choice(sp.nfree, 8, replace=False):
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            d = (S.face_terms(sp, faces, form, params, xp, None)[1] - S.face_terms(sp, faces, form, params, xm, None)[1])
            assert d / (2 * h) == pytest.approx(r[k], rel=1e-6, abs=1e-8 * np.abs(r).max())
        assert np.abs(A - A.T).max() < 1e-12 * np.abs(A).max()
