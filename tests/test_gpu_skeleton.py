"""Face integrals of the hexahedral operator on the device against the oracle (oracle/skeleton.py, pinned in
tests/test_oracle_skeleton.py), for Q1-Q3 on perturbed (non-affine) meshes with Dirichlet faces.  Skeleton
(interior facet) integrals:

* the sparsity pattern is the volume pattern plus the cross blocks of every facet (skeleton_keys!), bit for bit
  against a Python set of pairs; facets of one colour share no free dof (color_facets);
* residual!, Jv, the assembled CSC values and the energy equal the oracle's volume terms plus its skeleton terms
  (1e-12), for the built-in gradient-jump penalty and for a nonlinear facet form compiled from HIP source;
* the closed form of the penalty of a kinked field; a cell with rotated local axes (orientation 5) gives the
  operator of the unrotated mesh;
* facets that do not coincide, a (cell, face) used twice, and BlockJacobi on such an operator are refused.
Boundary groups with a captured integrand: the built-in follower pressure and a nonlinear form compiled from HIP
source, alone and together with a skeleton: residual!, Jv, the CSC values and the energy equal the oracle's (1e-12),
the pattern stays the volume pattern.
"""
import numpy as np
import pytest

from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk._lib import HFKError
from hfk.functionals import KernelFunctional
from hfk.mesh import dirichlet_closed_x0, dirichlet_none, hex_space, interior_facets
from oracle import skeleton as S
from skeleton_forms import NONLINEAR_BOUNDARY_SOURCE, NONLINEAR_FACET_SOURCE

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def make(hip_ctx, oracle, order, n=(3, 2, 2), law="neohookean", perturb=0.15, dirichlet=dirichlet_closed_x0):
    sp = hex_space(n, order, dirichlet=dirichlet, perturb=perturb, seed=order)
    u = lambda X: 0.05 * np.stack([np.sin(2 * X[:, 0] + X[:, 1]), X[:, 2] * X[:, 0], np.cos(X[:, 1])], axis=1)
    dv = sp.dirichlet_values(u)
    op = KernelOperator(HyperelasticForm(law, 2.0, 1.0, degree=2 * order), sp, jacobian="assembled", dirichlet_values=dv,
                        ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    return sp, dv, op, orc


def expected_pattern(sp, fac):
    pairs = set()
    for ids in sp.cell_dofs:
        f = [int(i) for i in ids if i > 0]
        pairs.update((c, r) for c in f for r in f)
    pairs |= S.skeleton_pairs(sp, fac)
    cols = sorted(pairs)
    colptr = np.ones(sp.nfree + 1, dtype=np.int64)
    for c, _ in cols:
        colptr[c] += 1
    colptr = np.cumsum(colptr) - np.arange(sp.nfree + 1)
    colptr[0] = 1
    return colptr, np.array([r for _, r in cols], dtype=np.int64)


def dense_csc(colptr, rowval, nz, n):
    import scipy.sparse as ssp

    return ssp.csc_matrix((nz, rowval - 1, colptr - 1), shape=(n, n)).toarray()


@pytest.mark.parametrize("order,form", [(1, "builtin"), (2, "builtin"), (3, "builtin"), (2, "compiled")])
def test_skeleton_against_oracle(hip_ctx, oracle, order, form):
    print(f"skeleton Q{order} {form}", flush=True)
    sp, dv, op, orc = make(hip_ctx, oracle, order)
    fac = interior_facets(sp)
    if form == "builtin":
        params, oform = [0.9], S.GradJump
        op.set_skeleton(fac, params=params)
    else:
        params, oform = [0.9, 2.0, 0.6, 1.5], S.NonlinearFacet
        op.set_skeleton(fac, source=NONLINEAR_FACET_SOURCE, name="NonlinearFacet", params=params)
    colptr, rowval, nzpos = op.pattern()
    c2, r2 = expected_pattern(sp, fac)
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2)
    colors = op.skeleton_colors()
    for c in range(1, colors.max() + 1):
        # two facets of one colour share no free dof of their cells (a facet's own two cells share their face's)
        per = [{int(i) for i in np.r_[sp.cell_dofs[fac[0][f]], sp.cell_dofs[fac[2][f]]] if i > 0}
               for f in np.nonzero(colors == c)[0]]
        assert sum(len(d) for d in per) == len(set().union(*per))
    rng = np.random.default_rng(order)
    x = 0.01 / order * rng.standard_normal(sp.nfree)  # small enough that det F > 0 at every point
    dx = rng.standard_normal(sp.nfree)
    rs, es, ys = S.skeleton_terms(sp, fac, oform, params, x, dv, dx=dx)
    b = op.residual(np.zeros(sp.nfree), x)
    assert rel(b, orc.residual(x) + rs) < 1e-12
    mf = MatrixFreeJacobian(op, "full64").prepare(x)
    prep = orc.prepare(x)
    assert rel(mf @ dx, orc.jv(prep, dx) + ys) < 1e-12
    A = op.jacobian(op.allocate_jacobian(), x)
    Ad = dense_csc(colptr, rowval, A.data, sp.nfree)
    Ao = orc.csc(x).toarray() + S.skeleton_matrix(sp, fac, oform, params, x, dv)
    assert np.abs(Ad - Ao).max() < 1e-12 * np.abs(Ao).max()
    e = KernelFunctional(op)(x)
    assert e == pytest.approx(orc.energy(x) + es, rel=1e-12)
    print(f"Q{order} {form}: {len(fac[0])} facets, {colors.max()} colours, nnz {rowval.size}", flush=True)


def test_kink_closed_form_on_device(hip_ctx, oracle):
    sp = hex_space((2, 3, 2), 2, dirichlet=dirichlet_none)
    op = KernelOperator(HyperelasticForm("neohookean", 2.0, 1.0, degree=4), sp, jacobian="matrixfree", ctx=hip_ctx)
    x = sp.free_values(lambda X: np.repeat(np.maximum(X[:, :1] - 0.5, 0.0), 3, axis=1))
    e0 = KernelFunctional(op)(x)
    op.set_skeleton(interior_facets(sp), params=[0.7])
    e1 = KernelFunctional(op)(x)
    assert e1 - e0 == pytest.approx(1.5 * 0.7 / np.sqrt(6), rel=1e-12)
    op.set_skeleton(None)
    assert KernelFunctional(op)(x) == e0


def test_rotated_cell(hip_ctx, oracle):
    from test_oracle_skeleton import rotated_pair

    sp, rot, o = rotated_pair(2)
    fac0 = interior_facets(sp)
    fac1 = (fac0[0], fac0[1], fac0[2], np.array([0], dtype=np.int32), np.array([o], dtype=np.int32))
    x = 0.005 * np.random.default_rng(1).standard_normal(sp.nfree)
    out = []
    for space, fac in ((sp, fac0), (rot, fac1)):
        op = KernelOperator(HyperelasticForm("neohookean", 2.0, 1.0, degree=4), space, jacobian="assembled", ctx=hip_ctx)
        op.set_skeleton(fac, source=NONLINEAR_FACET_SOURCE, name="NonlinearFacet", params=[0.9, 2.0, 0.6, 1.5])
        colptr, rowval, _ = op.pattern()
        A = op.jacobian(op.allocate_jacobian(), x)
        out.append((op.residual(np.zeros(space.nfree), x), dense_csc(colptr, rowval, A.data, space.nfree)))
    assert rel(out[1][0], out[0][0]) < 1e-12
    assert np.abs(out[1][1] - out[0][1]).max() < 1e-12 * np.abs(out[0][1]).max()


def test_refused(hip_ctx, oracle):
    sp, dv, op, orc = make(hip_ctx, oracle, 2, n=(2, 2, 1))
    pc, pf, mc, mf, ori = interior_facets(sp)
    with pytest.raises(HFKError):  # faces that do not coincide
        op.set_skeleton((pc, pf, mc, pf, ori))
    with pytest.raises(HFKError):  # a (cell, face) listed twice
        op.set_skeleton((np.r_[pc, pc[:1]], np.r_[pf, pf[:1]], np.r_[mc, mc[:1]], np.r_[mf, mf[:1]], np.r_[ori, ori[:1]]))
    with pytest.raises(HFKError):  # an orientation under which the faces do not match
        op.set_skeleton((pc, pf, mc, mf, np.full_like(ori, 1)))
    op.set_skeleton((pc, pf, mc, mf, ori))
    mfj = MatrixFreeJacobian(op, "full64").prepare(np.zeros(sp.nfree))
    from hfk.solvers import BlockJacobi

    with pytest.raises(HFKError):  # as the reference, src/preconditioners.jl:141
        BlockJacobi(mfj)


def some_boundary_faces(sp):
    from hfk.mesh import boundary_faces

    return tuple(np.concatenate(a) for a in zip(boundary_faces(sp, 0, 1), boundary_faces(sp, 1, 0),
                                                boundary_faces(sp, 2, 1)))


@pytest.mark.parametrize("order,form,skeleton", [(1, "follower", False), (2, "follower", False),
                                                 (2, "compiled", False), (3, "compiled", False),
                                                 (2, "compiled", True)])
def test_boundary_form_against_oracle(hip_ctx, oracle, order, form, skeleton):
    print(f"boundary form Q{order} {form} skeleton={skeleton}", flush=True)
    sp, dv, op, orc = make(hip_ctx, oracle, order)
    fc, fl = some_boundary_faces(sp)
    if form == "follower":
        params, oform = [2.0], S.FollowerPressure
        op.set_boundary_form(fc, fl, params=params)
    else:
        params, oform = [0.7, 3.0, 0.4, 1.3], S.NonlinearBoundary
        op.set_boundary_form(fc, fl, source=NONLINEAR_BOUNDARY_SOURCE, name="NonlinearBoundary", params=params)
    fac = interior_facets(sp)
    if skeleton:
        op.set_skeleton(fac, params=[0.9])
    colptr, rowval, _ = op.pattern()
    c0, r0, _ = orc.pattern()
    if not skeleton:  # a boundary group couples the dofs of one cell: the volume pattern
        assert np.array_equal(colptr, c0) and np.array_equal(rowval, r0)
    rng = np.random.default_rng(order)
    x = 0.01 / order * rng.standard_normal(sp.nfree)
    dx = rng.standard_normal(sp.nfree)
    rb, eb, yb = S.face_terms(sp, (fc, fl), oform, params, x, dv, dx=dx)
    Ab = S.face_matrix(sp, (fc, fl), oform, params, x, dv)
    if skeleton:
        rs, es, ys = S.face_terms(sp, fac, S.GradJump, [0.9], x, dv, dx=dx)
        rb, eb, yb = rb + rs, eb + es, yb + ys
        Ab = Ab + S.face_matrix(sp, fac, S.GradJump, [0.9], x, dv)
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x) + rb) < 1e-12
    mf = MatrixFreeJacobian(op, "full64").prepare(x)
    assert rel(mf @ dx, orc.jv(orc.prepare(x), dx) + yb) < 1e-12
    A = op.jacobian(op.allocate_jacobian(), x)
    Ad = dense_csc(colptr, rowval, A.data, sp.nfree)
    Ao = orc.csc(x).toarray() + Ab
    assert np.abs(Ad - Ao).max() < 1e-12 * np.abs(Ao).max()
    assert KernelFunctional(op)(x) == pytest.approx(orc.energy(x) + eb, rel=1e-12)
    with pytest.raises(HFKError):  # a face listed twice
        op.set_boundary_form(np.r_[fc, fc[:1]], np.r_[fl, fl[:1]])
    op.set_boundary_form(None, None)
    op.set_skeleton(None)
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < 1e-12
