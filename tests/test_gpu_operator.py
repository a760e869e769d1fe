"""GPU parity of libhfk against the oracle (the CPU restatement of the reference's FP64 algorithms).

Bars: integer outputs (colours, CSC pattern, nzpos) bit-exact; FP64 outputs (tangents, Jv, residual, CSR
values) within 1e-12 relative in the 2-norm (the reference's TOL, test/runtests.jl:9); FP32 tangent storage
within 1e-5 relative (reported, mixed precision).
"""
import numpy as np
import pytest

import hfk
from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk.mesh import hex_space
from hfk.synthetic import normal, smooth_field

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def problem(order, P, law, n, perturb=0.0, seed=0, amp=0.02):
    """Random dof values (amplitude amp/order: the gradients of a rough field grow with the order, and the
    neo-Hookean law needs det F > 0 at every point)."""
    amp = amp / order
    sp = hex_space(n, order, perturb=perturb, seed=seed)
    rng = np.random.default_rng(seed + 100)
    dv = amp * rng.standard_normal(sp.ndir)
    x = amp * rng.standard_normal(sp.nfree)
    return sp, dv, x


CASES = [
    (1, 2, "svk", (4, 3, 3), 0.0),
    (1, 3, "neohookean", (3, 3, 2), 0.1),
    (2, 3, "svk", (3, 2, 2), 0.0),
    (2, 3, "neohookean", (4, 3, 3), 0.15),
    (2, 3, "linear", (3, 3, 3), 0.1),
    (2, 4, "neohookean", (2, 2, 2), 0.1),
    (3, 4, "neohookean", (2, 2, 2), 0.1),
    (3, 4, "svk", (3, 2, 1), 0.0),
]


@pytest.mark.parametrize("order,P,law,n,perturb", CASES)
def test_tangents_jv_residual(oracle, hip_ctx, order, P, law, n, perturb):
    print(f"case order={order} P={P} law={law} n={n}", flush=True)
    sp, dv, x = problem(order, P, law, n, perturb)
    form = HyperelasticForm(law, 2.0, 1.0, degree=2 * (P - 1))
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    assert np.array_equal(op.colors(), orc.colors())
    # prepare: stored reference-coordinate tangents
    mf = MatrixFreeJacobian(op, "full64").prepare(x)
    prep = orc.prepare(x)
    Dref = oracle.reference_coordinate_tangents(*prep)
    assert np.isfinite(Dref).all()
    assert rel(mf.tangents(), Dref) < TOL
    # Jv
    dx = np.random.default_rng(7).standard_normal(sp.nfree)
    yref = orc.jv(prep, dx)
    assert rel(mf @ dx, yref) < TOL
    # residual
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < TOL


@pytest.mark.parametrize("storage,tol", [("sym64", TOL), ("full32", 1e-5), ("sym32", 1e-5), ("fac64", TOL),
                                         ("fac32", 1e-5)])
def test_storage_variants(oracle, hip_ctx, storage, tol):
    sp, dv, x = problem(2, 3, "neohookean", (4, 3, 3), 0.1)
    op = KernelOperator(HyperelasticForm("neohookean"), sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, dvals=dv)
    dx = np.random.default_rng(8).standard_normal(sp.nfree)
    yref = orc.jv(orc.prepare(x), dx)
    mf = MatrixFreeJacobian(op, storage).prepare(x)
    e = rel(mf @ dx, yref)
    print(f"storage {storage}: rel error {e:.3e}", flush=True)
    assert e < tol
    info = op.info()
    nent = {"sym": 45, "fac": 17, "ful": 81}[storage[:3]]
    assert info["tangent_bytes"] >= sp.ncells * 27 * nent * (4 if storage.endswith("32") else 8)


@pytest.mark.parametrize("order,P,law,n,perturb", [(2, 3, "neohookean", (4, 3, 3), 0.15), (2, 3, "linear", (3, 3, 2), 0.1),
                                                   (1, 2, "neohookean", (4, 3, 3), 0.1), (1, 3, "neohookean", (3, 3, 2), 0.1),
                                                   (3, 4, "neohookean", (2, 2, 2), 0.1), (2, 4, "linear", (2, 2, 2), 0.0)])
def test_factorised_storage(oracle, hip_ctx, order, P, law, n, perturb):
    """The factorised storage (17 factors per point, G d_ck + b H_Kk H_Mc + l H_Kc H_Mk): the 81 entries it
    represents equal the oracle's reference-coordinate tangents (1e-12), also with a shear-modulus field; Jv equals
    the oracle's (fac64 1e-12, fac32 1e-5), in the colour-ordered and the atomic single-launch products; the node
    blocks of BlockJacobi equal those built from the full storage.  St Venant-Kirchhoff and compiled laws, which
    have no such form, are refused."""
    from hfk.solvers import BlockJacobi

    print(f"factorised order={order} P={P} law={law} n={n}", flush=True)
    sp, dv, x = problem(order, P, law, n, perturb)
    form = HyperelasticForm(law, 2.0, 1.0, degree=2 * (P - 1))
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    mu = np.exp(0.4 * np.random.default_rng(5).standard_normal(sp.ncells * P ** 3))
    dx = np.random.default_rng(6).standard_normal(sp.nfree)
    for field in (False, True):
        if field:
            op.set_mu_field(mu)
            orc.set_mu_field(mu)
        prep = orc.prepare(x)
        yref = orc.jv(prep, dx)
        mf = MatrixFreeJacobian(op, "fac64").prepare(x)
        assert rel(mf.tangents(), oracle.reference_coordinate_tangents(*prep)) < TOL
        assert rel(mf @ dx, yref) < TOL
        op.set_deterministic(False)
        assert rel(mf @ dx, yref) < TOL
        op.set_deterministic(True)
        b64 = BlockJacobi(mf).blocks()[1]
        bfull = BlockJacobi(MatrixFreeJacobian(op, "full64").prepare(x)).blocks()[1]
        assert rel(b64, bfull) < TOL
        assert rel(MatrixFreeJacobian(op, "fac32").prepare(x) @ dx, yref) < 1e-5
    assert op.info()["tangent_bytes"] >= sp.ncells * P ** 3 * 17 * 4


@pytest.mark.parametrize("order,P,law,n", [(2, 3, "neohookean", (4, 3, 3)), (2, 3, "linear", (3, 3, 2)),
                                           (1, 2, "neohookean", (4, 3, 3)), (3, 4, "neohookean", (2, 2, 2)),
                                           (1, 3, "neohookean", (3, 2, 2))])
def test_affine_factorised_storage(oracle, hip_ctx, order, P, law, n):
    """The affine factorised storage (11 numbers per point, the metric and det J per cell) on parallelepiped cells --
    the box mesh under a global affine map, so that iJt is a full matrix: the 81 entries it represents equal the
    oracle's tangents (1e-12), also with a shear-modulus field; Jv equals the oracle's (afac64 1e-12, afac32 1e-5) in
    both product modes; BlockJacobi's blocks equal those of the full storage.  Non-affine cells are refused."""
    from hfk.solvers import BlockJacobi

    print(f"affine factorised order={order} P={P} law={law} n={n}", flush=True)
    sp, dv, x = problem(order, P, law, n, 0.0)
    A = np.array([[1.1, 0.2, -0.1], [0.05, 0.9, 0.3], [-0.2, 0.1, 1.2]])
    sp.coords[:] = sp.coords @ A.T + np.array([0.3, -0.1, 0.2])
    form = HyperelasticForm(law, 2.0, 1.0, degree=2 * (P - 1))
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    mu = np.exp(0.4 * np.random.default_rng(5).standard_normal(sp.ncells * P ** 3))
    dx = np.random.default_rng(6).standard_normal(sp.nfree)
    for field in (False, True):
        if field:
            op.set_mu_field(mu)
            orc.set_mu_field(mu)
        prep = orc.prepare(x)
        yref = orc.jv(prep, dx)
        mf = MatrixFreeJacobian(op, "afac64").prepare(x)
        assert rel(mf.tangents(), oracle.reference_coordinate_tangents(*prep)) < TOL
        assert rel(mf @ dx, yref) < TOL
        op.set_deterministic(False)
        assert rel(mf @ dx, yref) < TOL
        op.set_deterministic(True)
        bfac = BlockJacobi(mf).blocks()[1]
        bfull = BlockJacobi(MatrixFreeJacobian(op, "full64").prepare(x)).blocks()[1]
        assert rel(bfac, bfull) < TOL
        assert rel(MatrixFreeJacobian(op, "afac32").prepare(x) @ dx, yref) < 1e-5
    assert op.info()["tangent_bytes"] >= sp.ncells * P ** 3 * 11 * 4
    spp, dvp, xp = problem(order, P, law, n, 0.1)  # trilinear (non-affine) cells
    opp = KernelOperator(form, spp, jacobian="matrixfree", dirichlet_values=dvp, ctx=hip_ctx)
    with pytest.raises(hfk._lib.HFKError, match="ArgumentError"):
        MatrixFreeJacobian(opp, "afac64").prepare(xp)


def test_factorised_storage_refused(hip_ctx):
    from hfk.laws import SOURCES, CompiledLaw

    sp, dv, x = problem(2, 3, "svk", (2, 2, 2))
    op = KernelOperator(HyperelasticForm("svk"), sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    with pytest.raises(hfk._lib.HFKError, match="ArgumentError"):
        MatrixFreeJacobian(op, "fac64").prepare(x)
    with pytest.raises(hfk._lib.HFKError, match="ArgumentError"):
        MatrixFreeJacobian(op, "afac64").prepare(x)
    op2 = KernelOperator(HyperelasticForm("neohookean"), sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    src, name = SOURCES["neohookean"]
    op2.set_law(CompiledLaw(src, name, ctx=hip_ctx))
    with pytest.raises(hfk._lib.HFKError, match="ArgumentError"):
        MatrixFreeJacobian(op2, "fac32").prepare(x)


def test_mu_field(oracle, hip_ctx):
    sp, dv, x = problem(2, 3, "neohookean", (3, 3, 2), 0.05)
    mu = np.exp(0.5 * np.random.default_rng(3).standard_normal(sp.ncells * 27))
    op = KernelOperator(HyperelasticForm("neohookean"), sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    op.set_mu_field(mu)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, dvals=dv)
    orc.set_mu_field(mu)
    dx = np.random.default_rng(9).standard_normal(sp.nfree)
    prep = orc.prepare(x)
    mf = MatrixFreeJacobian(op).prepare(x)
    assert rel(mf @ dx, orc.jv(prep, dx)) < TOL
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < TOL


# (order, points per direction): the coarse Q1 levels of the p-multigrid keep the fine rule (3 points under Q2,
# 4 under Q3)
@pytest.mark.parametrize("order,P,law,n", [(1, 2, "neohookean", (4, 3, 3)), (1, 3, "neohookean", (3, 3, 2)),
                                           (1, 4, "svk", (3, 2, 2)), (2, 3, "svk", (3, 2, 2)),
                                           (2, 3, "neohookean", (4, 4, 3)), (3, 4, "neohookean", (2, 2, 1))])
def test_pattern_and_assembly(oracle, hip_ctx, order, P, law, n):
    print(f"assembly order={order} P={P} law={law} n={n}", flush=True)
    sp, dv, x = problem(order, P, law, n, 0.1)
    op = KernelOperator(HyperelasticForm(law, degree=2 * P - 2), sp, dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = orc.pattern()
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    A = op.jacobian(op.allocate_jacobian(), x)
    ref = orc.jacobian(x, p2, r2.size)
    assert rel(A.data, ref) < TOL
    # matrix-free = assembled . dx (test/test_matrixfree.jl:74-96)
    mf = MatrixFreeJacobian(op).prepare(x)
    dx = np.random.default_rng(4).standard_normal(sp.nfree)
    assert rel(mf @ dx, A @ dx) < 1e-11


@pytest.mark.parametrize("nfree_pad", [0, 1])
def test_pattern_random_numbering(oracle, hip_ctx, nfree_pad):
    """The sparsity pattern and positions (built on the device by a radix sort of (column, row) keys) equal the
    oracle's for a randomly renumbered dof table: no locality in the keys.  With nfree_pad, the free ids are spread
    over the next power of two (columns that no cell holds stay empty; the column bits of the keys are then full)."""
    import dataclasses

    sp, dv, x = problem(2, 3, "svk", (5, 4, 3), 0.1)
    rng = np.random.default_rng(7 + nfree_pad)
    n = sp.nfree if not nfree_pad else 1 << int(np.ceil(np.log2(sp.nfree + 1)))
    newid = rng.permutation(n)[: sp.nfree] + 1  # new id of old free id k: newid[k - 1]
    cd = sp.cell_dofs.copy()
    m = cd > 0
    cd[m] = newid[cd[m] - 1]
    sp2 = dataclasses.replace(sp, cell_dofs=np.ascontiguousarray(cd, dtype=np.int32), nfree=n)
    print(f"pattern numbering nfree={n} cells={cd.shape[0]}", flush=True)
    op = KernelOperator(HyperelasticForm("svk"), sp2, dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp2, "svk", 2.0, 1.0, dvals=dv)
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = orc.pattern()
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    if nfree_pad:
        assert (np.diff(colptr) == 0).sum() == n - sp.nfree  # the empty columns
    xp = np.zeros(n)
    xp[newid - 1] = x
    A = op.jacobian(op.allocate_jacobian(), xp)
    assert rel(A.data, orc.jacobian(xp, p2, r2.size)) < TOL


@pytest.mark.parametrize("order,P,storage", [(2, 3, "full64"), (2, 3, "sym32"), (1, 2, "full64"), (3, 4, "sym64")])
def test_atomic_single_launch(oracle, hip_ctx, order, P, storage):
    """The single-launch product with atomic adds equals the oracle to the same bar (its summation order differs
    from the colour order, so only to rounding)."""
    sp, dv, x = problem(order, P, "neohookean", (5, 4, 3) if order < 3 else (3, 2, 2), 0.1)
    op = KernelOperator(HyperelasticForm("neohookean", degree=2 * (P - 1)), sp, jacobian="matrixfree",
                        dirichlet_values=dv, ctx=hip_ctx, deterministic=False)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, points1d=P, dvals=dv)
    dx = np.random.default_rng(17).standard_normal(sp.nfree)
    yref = orc.jv(orc.prepare(x), dx)
    mf = MatrixFreeJacobian(op, storage).prepare(x)
    tol = TOL if storage.endswith("64") else 1e-5
    y1 = mf @ dx
    assert rel(y1, yref) < tol
    op.set_deterministic(True)
    y2 = mf @ dx
    assert rel(y1, y2) < (1e-14 if storage.endswith("64") else 1e-12)


def test_errors(hip_ctx):
    sp, dv, x = problem(2, 3, "svk", (2, 2, 2))
    op = KernelOperator(HyperelasticForm("svk"), sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    mf = MatrixFreeJacobian(op)
    with pytest.raises(hfk._lib.HFKError, match="ErrorException"):
        mf @ np.zeros(sp.nfree)  # mul! before prepare!
    with pytest.raises(ValueError):
        mf.prepare(np.zeros(sp.nfree + 1))
    with pytest.raises(ValueError):
        MatrixFreeJacobian(op, "bf16")


def test_benchmark_size_properties(oracle, hip_ctx):
    """64^3 Q2 neo-Hookean at the smooth random state: J is symmetric (w.Jv = v.Jw), linear, and equals
    the oracle's product on the same state to 1e-12 (oracle on the host cores)."""
    sp = hex_space((64, 64, 64), 2)
    u = smooth_field(2)
    x, dvals = sp.free_values(u), sp.dirichlet_values(u)
    op = KernelOperator(HyperelasticForm("neohookean"), sp, jacobian="matrixfree", dirichlet_values=dvals, ctx=hip_ctx)
    mf = MatrixFreeJacobian(op).prepare(x)
    v = normal(3, sp.nfree)
    w = normal(5, sp.nfree)
    Jv, Jw = mf @ v, mf @ w
    assert abs(w @ Jv - v @ Jw) / abs(w @ Jv) < 1e-11
    assert rel(mf @ (2.0 * v - 3.0 * w), 2.0 * Jv - 3.0 * Jw) < 1e-13
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, dvals=dvals)
    assert rel(Jv, orc.jv(orc.prepare(x), v)) < TOL


def test_newton_quickstart_matches_oracle(oracle, hip_ctx):
    """The reference's quick start (README.md:61-83) on a 10x5x5 mesh: Newton from x = 0 with host LU on the
    device residual and CSC Jacobian follows the oracle's Newton iterate by iterate, to 1e-10 of the residual."""
    import scipy.sparse.linalg as spla

    from hfk.mesh import dirichlet_open_x_faces

    sp = hex_space((10, 5, 5), 2, dirichlet=dirichlet_open_x_faces)
    dv = sp.dirichlet_values(lambda X: np.stack([0.05 * X[:, 0], 0.0 * X[:, 0], 0.0 * X[:, 0]], axis=1))
    op = KernelOperator(HyperelasticForm("svk", 2.0, 1.0, degree=4), sp, dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=dv)
    A = op.allocate_jacobian()
    x = np.zeros(sp.nfree)
    xo = np.zeros(sp.nfree)
    b = op.residual(np.zeros(sp.nfree), x)
    bo = orc.residual(xo)
    r0 = np.linalg.norm(bo)
    assert rel(b, bo) < TOL
    for k in range(12):
        op.jacobian(A, x)
        x += spla.splu(A.tocsc()).solve(-b)
        xo += spla.splu(orc.csc(xo).tocsc()).solve(-bo)
        op.residual(b, x)
        bo = orc.residual(xo)
        print(f"newton {k}: |R|/|R0| gpu {np.linalg.norm(b) / r0:.3e} oracle {np.linalg.norm(bo) / r0:.3e}", flush=True)
        assert np.linalg.norm(b - bo) <= 1e-10 * r0
        assert rel(x, xo) < 1e-10
        if np.linalg.norm(bo) <= 1e-10 * r0:
            break
    assert np.linalg.norm(b) <= 1e-10 * r0 and k <= 8


@pytest.mark.parametrize("order,P,n", [(2, 3, (1, 1, 1)), (2, 3, (5, 1, 1)), (2, 3, (3, 3, 1)), (3, 4, (1, 1, 2)),
                                       (1, 2, (1, 1, 1))])
def test_small_and_ragged_meshes(oracle, hip_ctx, order, P, n):
    """A single cell, ragged meshes (the last batch partly padding) and a mesh that fills its batch exactly: tangents,
    Jv in every storage, residual, assembled values and the energy equal the oracle's."""
    from hfk.functionals import KernelFunctional

    print(f"small mesh order={order} n={n}", flush=True)
    sp, dv, x = problem(order, P, "neohookean", n, perturb=0.0)
    form = HyperelasticForm("neohookean", 2.0, 1.0, degree=2 * (P - 1))
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, points1d=P, dvals=dv)
    prep = orc.prepare(x)
    dx = np.random.default_rng(9).standard_normal(sp.nfree)
    yref = orc.jv(prep, dx)
    for st, tol in (("full64", TOL), ("sym64", TOL), ("full32", 1e-5), ("sym32", 1e-5), ("fac64", TOL), ("fac32", 1e-5)):
        assert rel(MatrixFreeJacobian(op, st).prepare(x) @ dx, yref) < tol
    assert rel(MatrixFreeJacobian(op, "full64").prepare(x).tangents(), oracle.reference_coordinate_tangents(*prep)) < TOL
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < TOL
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = orc.pattern()
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    A = op.jacobian(op.allocate_jacobian(), x)
    assert rel(A.data, orc.jacobian(x, p2, r2.size)) < TOL
    assert KernelFunctional(op)(x) == pytest.approx(orc.energy(x), rel=1e-12)


def test_check_against_reference(oracle, hip_ctx):
    """check (src/verify.jl:9-17) with the oracle as the reference assembly: both differences below 1e-12, for the
    assembled and the matrix-free kinds; a wrong reference (scaled residual, shifted matrix) is reported."""
    from hfk.verify import check

    sp, dv, x = problem(2, 3, "svk", (3, 2, 2), 0.1)
    orc = oracle.OracleOp(sp, "svk", 2.0, 1.0, points1d=3, dvals=dv)
    for kind in ("assembled", "matrixfree"):
        op = KernelOperator(HyperelasticForm("svk", 2.0, 1.0, degree=4), sp, jacobian=kind, dirichlet_values=dv,
                            ctx=hip_ctx)
        r = check(op, orc.residual, orc.csc, x)
        assert r.res < TOL and r.jac < TOL
    bad = check(op, lambda v: 1.1 * orc.residual(v), lambda v: orc.csc(v) * 0.9, x)
    assert bad.res == pytest.approx(0.1 / 1.1, rel=1e-6) and bad.jac == pytest.approx(0.1 / 0.9, rel=1e-6)
