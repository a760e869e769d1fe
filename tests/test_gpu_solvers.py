"""GPU parity of the solver layer against the oracle and the reference's own solver tests.

* BlockJacobi inverted blocks = the oracle's (same first-seen block numbering), 1e-12;
* Arnoldi interval = the oracle's Arnoldi on the assembled J M^-1 from the same start vector;
* Chebyshev smooth! = the oracle's recurrence with the same interval, 1e-12;
* transfers: prolongation = the oracle's embedding matrix, restriction = its transpose, injection of a
  prolongated field is the identity (test/solvers/test_transfers.jl);
* p-multigrid FCG: same solution as a direct solve (<= 1e-7, test_pmultigrid.jl:37-51), at most 30
  iterations, nearly independent of h; a NaN state makes the solve return not converged at once
  (test_pmultigrid.jl:117-133).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk.mesh import dirichlet_closed_x0, hex_space
from hfk.solvers import (BlockJacobi, Chebyshev, ChebyshevState, MatrixFreeSolver, PMultigrid, Prolongation,
                         eigenvalue_bounds)

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def setup(hip_ctx, oracle, n=(3, 2, 2), order=2, law="neohookean", seed=0, amp=0.01):
    sp = hex_space(n, order, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    x = amp / order * rng.standard_normal(sp.nfree)
    dv = amp / order * rng.standard_normal(sp.ndir)
    form = HyperelasticForm(law, degree=2 * order)
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    mf = MatrixFreeJacobian(op).prepare(x)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    return sp, op, mf, orc, x


def test_block_jacobi(hip_ctx, oracle):
    sp, op, mf, orc, x = setup(hip_ctx, oracle)
    bj = BlockJacobi(mf)
    assert bj.ok
    dofs, Binv = bj.blocks()
    d2, B2 = orc.blocks_from_tangents(orc.prepare(x))
    assert np.array_equal(dofs, d2)
    assert rel(Binv, B2) < 1e-12
    r = np.random.default_rng(1).standard_normal(sp.nfree)
    y = bj.ldiv(np.zeros(sp.nfree), r)
    assert rel(y, oracle.block_inverse_matrix(d2, B2, sp.nfree) @ r) < 1e-12


@pytest.mark.parametrize("storage,perturb", [("fac64", 0.1), ("afac64", 0.0), ("fac32", 0.1), ("afac32", 0.0)])
def test_block_jacobi_factorised_storages(hip_ctx, oracle, storage, perturb):
    """The factorised storages form the nodal blocks from their factors in closed form ((g'Gg) I + (b + l) u u');
    their inverses equal the oracle's blocks from the 81-entry tangents (FP64 factors to 1e-12, FP32 to 1e-5)."""
    sp = hex_space((3, 2, 2), 2, dirichlet=dirichlet_closed_x0, perturb=perturb, seed=0)
    rng = np.random.default_rng(0)
    x = 0.005 * rng.standard_normal(sp.nfree)
    dv = 0.005 * rng.standard_normal(sp.ndir)
    op = KernelOperator(HyperelasticForm("neohookean", degree=4), sp, jacobian="matrixfree", dirichlet_values=dv,
                        ctx=hip_ctx)
    mf = MatrixFreeJacobian(op, storage).prepare(x)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, dvals=dv)
    bj = BlockJacobi(mf)
    assert bj.ok
    dofs, Binv = bj.blocks()
    d2, B2 = orc.blocks_from_tangents(orc.prepare(x))
    assert np.array_equal(dofs, d2)
    assert rel(Binv, B2) < (1e-12 if storage.endswith("64") else 1e-5)


def test_eigenvalue_bounds_and_smoother(hip_ctx, oracle):
    sp, op, mf, orc, x = setup(hip_ctx, oracle, n=(2, 2, 2))
    A = orc.csc(x)
    d2, B2 = orc.blocks_from_tangents(orc.prepare(x))
    Minv = oracle.block_inverse_matrix(d2, B2, sp.nfree)
    lo_r, hi_r = oracle.arnoldi_bounds(lambda v: A @ (Minv @ v), oracle.chebyshev_start(sp.nfree), 10)
    st = ChebyshevState(Chebyshev(degree=4), mf)
    assert st.ok and not st.indefinite
    assert st.lmax == pytest.approx(hi_r, rel=1e-8) and st.lmin == pytest.approx(lo_r, rel=1e-6)
    assert st.b == pytest.approx(1.1 * st.lmax, rel=1e-15) and st.a == pytest.approx(st.b / 20, rel=1e-15)
    lo, hi = eigenvalue_bounds(mf, 10)
    assert hi == pytest.approx(hi_r, rel=1e-8)
    rng = np.random.default_rng(5)
    rhs = rng.standard_normal(sp.nfree)
    x0 = rng.standard_normal(sp.nfree)
    for zero_guess in (True, False):
        res = np.zeros(sp.nfree)
        xg = st.smooth(x0.copy(), rhs, zero_guess=zero_guess, residual=res)
        xr, rr = oracle.chebyshev_smooth(A, Minv, x0, rhs, st.a, st.b, 4, zero_guess=zero_guess, want_residual=True)
        assert rel(xg, xr) < 1e-12 and rel(res, rr) < 1e-11
        xg2 = st.smooth(x0.copy(), rhs, zero_guess=zero_guess)
        assert rel(xg2, xr) < 1e-12


@pytest.mark.parametrize("qc,qf", [(1, 2), (2, 3), (1, 3)])
def test_transfers(hip_ctx, oracle, qc, qf):
    n = (3, 2, 2) if qf < 3 else (2, 2, 1)
    fs = hex_space(n, qf, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=2)
    cs = hex_space(n, qc, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=2)
    fop = KernelOperator(HyperelasticForm("svk", degree=2 * qf), fs, jacobian="matrixfree", ctx=hip_ctx)
    cop = KernelOperator(HyperelasticForm("svk", degree=2 * qf), cs, jacobian="matrixfree", ctx=hip_ctx)
    P = Prolongation(cop, fop)
    Pm = oracle.prolongation_matrix(cs, fs)
    rng = np.random.default_rng(3)
    xc = rng.standard_normal(cs.nfree)
    yf = rng.standard_normal(fs.nfree)
    assert rel(P @ xc, Pm @ xc) < 1e-14
    yc = P.restrict(np.zeros(cs.nfree), yf)
    assert rel(yc, Pm.T @ yf) < 1e-13
    if qc == 1:  # nested nodes: the injection of a prolongated field gives it back
        xi, _ = P.inject(P @ xc, np.zeros(fs.ndir))
        assert rel(xi, xc) < 1e-14


@pytest.mark.parametrize("law", ["linear", "neohookean"])
def test_galerkin_coarse_operator(hip_ctx, law):
    """The assembled coarse level keeps the working quadrature (src/pmultigrid.jl:248-251), so for a linear law it
    is the Galerkin product P' A P of the working operator to round-off; at a state of the neo-Hookean law (the
    coarse tangent taken at the injected state) the gap is O(h) (test/solvers/test_pmultigrid_levels.jl:40-63).
    The state is a quadratic field, which Q2 holds exactly and whose nodal Q1 interpolant has an O(h) gradient
    error."""
    def u(X):
        return 0.05 * np.stack([X[:, 1] * X[:, 2], X[:, 0] ** 2, X[:, 0] * X[:, 2]], axis=1)

    gaps = []
    for n in ((4, 4, 4), (8, 8, 8)):
        print(f"galerkin law={law} n={n}", flush=True)
        fs = hex_space(n, 2, dirichlet=dirichlet_closed_x0)
        cs = hex_space(n, 1, dirichlet=dirichlet_closed_x0)
        fop = KernelOperator(HyperelasticForm(law, degree=4), fs, jacobian="matrixfree",
                             dirichlet_values=fs.dirichlet_values(u), ctx=hip_ctx)
        cop = KernelOperator(HyperelasticForm(law, degree=4), cs, jacobian="assembled", ctx=hip_ctx)
        mf = MatrixFreeJacobian(fop).prepare(fs.free_values(u))
        pmg = PMultigrid([fop, cop], Chebyshev(degree=2), coarse_rtol=1e-3).update()
        assert pmg.ok
        P = Prolongation(cop, fop)
        rng = np.random.default_rng(21)
        gap = 0.0
        for _ in range(3):
            xc = rng.standard_normal(cop.nfree)
            ptap = P.restrict(np.zeros(cop.nfree), mf @ (P @ xc))
            gap = max(gap, rel(pmg.coarse_matvec(xc), ptap))
        gaps.append(gap)
    print("galerkin gaps", law, gaps, flush=True)
    if law == "linear":
        assert max(gaps) < 1e-12
    else:
        assert 1e-8 < gaps[1] < 0.6 * gaps[0]  # halving h about halves it


def make_levels(hip_ctx, n, law="neohookean", seed=0, amp=0.01):
    """Q2 working level and assembled Q1 coarse level on n^3 cells clamped on the closed face x = 0, at the
    smooth random state of amplitude amp (1e-2 keeps |grad u| <= 0.19: a positive definite tangent)."""
    from hfk.synthetic import smooth_field

    fs = hex_space(n, 2, dirichlet=dirichlet_closed_x0)
    cs = hex_space(n, 1, dirichlet=dirichlet_closed_x0)
    u = smooth_field(seed + 2, amp=amp)
    x = fs.free_values(u)
    fop = KernelOperator(HyperelasticForm(law, degree=4), fs, jacobian="matrixfree",
                         dirichlet_values=fs.dirichlet_values(u), ctx=hip_ctx)
    # the coarse level keeps the fine quadrature (src/pmultigrid.jl:248-251)
    cop = KernelOperator(HyperelasticForm(law, degree=4), cs, jacobian="assembled", ctx=hip_ctx)
    mf = MatrixFreeJacobian(fop).prepare(x)
    return fs, fop, cop, mf, x


def test_pmultigrid_solution_equals_direct(hip_ctx, oracle):
    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3).update()
    assert pmg.ok
    # the coarse solver's node-block ELL product = the scalar CSC product of the same matrix
    xc = np.random.default_rng(10).standard_normal(cop.nfree)
    assert rel(pmg.coarse_matvec(xc), pmg.coarse_matvec(xc, scalar_csc=True)) < 1e-14
    b = np.random.default_rng(11).standard_normal(fs.nfree)
    sol = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg)
    xs = np.zeros(fs.nfree)
    res = sol.solve(xs, mf, b)
    print("pmg fcg:", res, pmg.info(), flush=True)
    assert res.converged and res.iterations <= 30
    from hfk.synthetic import smooth_field

    # the oracle's Jacobian at the same state and Dirichlet values
    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    A = orc.csc(x)
    xd = spla.spsolve(A.tocsc(), b)
    assert rel(xs, xd) < 1e-7
    # CG with the nodal blocks alone also reaches the solution (more iterations)
    BlockJacobi(mf)
    xs2 = np.zeros(fs.nfree)
    res2 = MatrixFreeSolver("cg", rtol=1e-10, maxiter=2000, preconditioner="blockjacobi").solve(xs2, mf, b)
    assert res2.converged and rel(xs2, xd) < 1e-7
    assert res2.iterations > res.iterations


def test_pmultigrid_coarse_every(hip_ctx, oracle):
    """coarse_every = 3 (test_pmultigrid_levels.jl:98-107): four updates at changing states build the coarse level
    at updates 1 and 4; between them the coarse matrix is the lagged one, and FCG with the lagged coarse level
    still reaches the direct solution at the new state (test_pmultigrid.jl:107-114)."""
    from hfk.synthetic import smooth_field

    with pytest.raises(ValueError):
        PMultigrid([None, None], coarse_every=0)
    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3, coarse_every=3)
    # the every-update hierarchy gets its own coarse operator: a coarse operator's node blocks belong to the
    # hierarchy that assembles it
    cop1 = KernelOperator(HyperelasticForm("neohookean", degree=4), hex_space((4, 4, 4), 1, dirichlet=dirichlet_closed_x0),
                          jacobian="assembled", ctx=hip_ctx)
    every1 = PMultigrid([fop, cop1], Chebyshev(degree=4), coarse_rtol=1e-3)
    xc = np.random.default_rng(10).standard_normal(cop.nfree)
    b = np.random.default_rng(11).standard_normal(fs.nfree)
    dvals = fs.dirichlet_values(smooth_field(2, amp=0.01))
    ac = []
    for k in range(4):
        xk = (1.0 + 0.2 * k) * x
        mf.prepare(xk)
        pmg.update()
        assert pmg.ok
        ac.append(pmg.coarse_matvec(xc))
        assert pmg.coarse_builds() == (1 if k < 3 else 2)
        if k == 2:
            # the lagged coarse level of update 1 preconditions the tangent at this state
            xs = np.zeros(fs.nfree)
            res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
            assert res.converged and res.iterations <= 30
            orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=dvals)
            assert rel(xs, spla.spsolve(orc.csc(xk).tocsc(), b)) < 1e-7
            every1.update()
            assert every1.coarse_builds() == 1
            fresh = every1.coarse_matvec(xc)
    assert np.array_equal(ac[0], ac[1]) and np.array_equal(ac[0], ac[2])
    assert rel(ac[2], fresh) > 1e-6 and rel(ac[3], ac[2]) > 1e-6
    # at update 4 the rebuilt coarse matrix is the one an every-update hierarchy assembles at that state
    every1.update()
    assert every1.coarse_builds() == 2 and rel(ac[3], every1.coarse_matvec(xc)) < 1e-14


def test_pmultigrid_with_robin_faces(hip_ctx, oracle):
    """Springs k int u.v dGamma on the face x = 1 of both levels (each level's operator carries its own Robin term,
    the coarse one in its assembled matrix and node blocks): FCG with the p-multigrid reaches the direct solution of
    the oracle's Jacobian plus the oracle's face mass matrix."""
    from hfk.mesh import boundary_faces
    from hfk.synthetic import smooth_field

    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    cs = hex_space((4, 4, 4), 1, dirichlet=dirichlet_closed_x0)
    fc, fl = boundary_faces(fs, 0, 1)
    fop.set_robin(fc, fl, 5.0)
    cop.set_robin(*boundary_faces(cs, 0, 1), 5.0)
    mf.prepare(x)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3).update()
    assert pmg.ok
    b = np.random.default_rng(13).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
    print("pmg fcg with robin:", res, flush=True)
    assert res.converged and res.iterations <= 30
    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    A = orc.csc(x).toarray() + orc.face_mass(fc, fl, 5.0)[:fs.nfree, :fs.nfree]
    assert rel(xs, np.linalg.solve(A, b)) < 1e-7


def test_pmultigrid_fp32_coarse_values(hip_ctx, oracle):
    """The coarse PCG's products on FP32 copies of the coarse values (FP64 vectors and sums): the flexible outer
    solve still reaches the direct solution of the FP64 system (the coarse level only preconditions)."""
    from hfk.synthetic import smooth_field

    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3, coarse_precision="fp32").update()
    assert pmg.ok
    b = np.random.default_rng(11).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
    print("pmg fcg, fp32 coarse values:", res, pmg.info(), flush=True)
    assert res.converged and res.iterations <= 30
    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    assert rel(xs, spla.spsolve(orc.csc(x).tocsc(), b)) < 1e-7
    with pytest.raises(ValueError):
        PMultigrid([fop, cop], coarse_precision="bf16")


def test_pmultigrid_q3_lognormal(hip_ctx, oracle):
    """Q3 working level over an assembled Q1 coarse level with the fine 4-point rule, heterogeneous (lognormal)
    shear modulus, FP32 smoothers: the flexible CG reaches the direct solution."""
    from hfk.synthetic import lognormal_field, smooth_field

    n = (3, 3, 3)
    fs = hex_space(n, 3, dirichlet=dirichlet_closed_x0)
    cs = hex_space(n, 1, dirichlet=dirichlet_closed_x0)
    u = smooth_field(2, amp=0.01)
    x = fs.free_values(u)
    form = HyperelasticForm("neohookean", degree=6)
    assert form.points1d == 4
    fop = KernelOperator(form, fs, jacobian="matrixfree", dirichlet_values=fs.dirichlet_values(u), ctx=hip_ctx)
    cop = KernelOperator(form, cs, jacobian="assembled", ctx=hip_ctx)
    mup = lognormal_field(4, sigma=1.0, ell=0.3)(fs.quadrature_points(4).reshape(-1, 3))
    fop.set_mu_field(mup)
    cop.set_mu_field(mup)
    mf = MatrixFreeJacobian(fop).prepare(x)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-2, smoother_storage="sym32").update()
    assert pmg.ok
    xc = np.random.default_rng(10).standard_normal(cs.nfree)
    assert rel(pmg.coarse_matvec(xc), pmg.coarse_matvec(xc, scalar_csc=True)) < 1e-14
    b = np.random.default_rng(12).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
    print("q3 pmg fcg:", res, flush=True)
    assert res.converged and res.iterations <= 30
    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(u))
    orc.set_mu_field(mup)
    xd = spla.spsolve(orc.csc(x).tocsc(), b)
    assert rel(xs, xd) < 1e-7


@pytest.mark.parametrize("restart", [50, 10])
def test_gmres_blockjacobi_matches_oracle(hip_ctx, oracle, restart):
    """GMRES (the reference's default MatrixFreeSolver method) right-preconditioned by the nodal blocks: the same
    iteration count as the oracle's restatement on the assembled matrix (+-1 for rounding at the tolerance), the
    direct solution to 1e-7, restarts included (GMRES(10) needs ~400 iterations here, GMRES(50) ~220)."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (3, 3, 3))
    from hfk.synthetic import smooth_field

    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    A = orc.csc(x).tocsc()
    bj = BlockJacobi(mf)
    dofs, Binv = bj.blocks()
    Minv = oracle.block_inverse_matrix(dofs, Binv, fs.nfree)
    b = np.random.default_rng(13).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("gmres", rtol=1e-9, maxiter=1000, restart=restart, preconditioner="blockjacobi").solve(xs, mf, b)
    xo, ito, oko = oracle.gmres_right(A, b, lambda v: Minv @ v, 1e-9, restart, 1000)
    print(f"gmres(restart={restart}): device {res}, oracle {ito} iterations", flush=True)
    assert res.converged and oko
    assert abs(res.iterations - ito) <= 1
    xd = spla.spsolve(A, b)
    assert rel(xs, xd) < 1e-7 and rel(xs, xo) < 1e-7
    assert res.residual <= 1e-9 * res.rhs_norm * 1.01


@pytest.mark.parametrize("krylov", ["fgmres", "gmres"])
def test_gmres_pmultigrid(hip_ctx, oracle, krylov):
    """FGMRES with one p-multigrid V-cycle as right preconditioner reaches the direct solution.  Plain GMRES
    needs a fixed preconditioner: with the coarse level solved to 1e-12 it does too (with an inexact coarse solve
    the cycle varies, its residual estimate no longer equals the true residual, and the flexible method is the
    one to use, as src/pmultigrid.jl:237-239 says)."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3 if krylov == "fgmres" else 1e-12,
                     coarse_maxiter=2000).update()
    b = np.random.default_rng(11).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver(krylov, rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
    print(f"{krylov} + pmg: {res}", flush=True)
    assert res.converged and res.iterations <= 30
    from hfk.synthetic import smooth_field

    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    assert rel(xs, spla.spsolve(orc.csc(x).tocsc(), b)) < 1e-7


@pytest.mark.parametrize("coarse_rtol", [1e-2, 0.3])
def test_gmres_default_pmultigrid_converged_means_true_residual(hip_ctx, coarse_rtol):
    """Plain GMRES with the default p-multigrid (inexact coarse PCG: a cycle that varies between applications).
    Its residual estimate is then not the true residual, so a converged status must rest on the true residual:
    converged implies ||b - A x|| <= rtol ||b||, and the solve keeps restarting until that holds."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=coarse_rtol).update()
    b = np.random.default_rng(12).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("gmres", rtol=1e-10, maxiter=200, preconditioner=pmg).solve(xs, mf, b)
    print(f"gmres + default pmg (coarse_rtol {coarse_rtol}): {res}", flush=True)
    true = np.linalg.norm(b - mf @ xs)
    assert res.converged
    assert res.residual == pytest.approx(true, rel=1e-6)
    assert true <= 1e-10 * np.linalg.norm(b)


def test_gmres_failed_preconditioner_returns_at_once(hip_ctx):
    """A p-multigrid that is not ok (NaN state) makes GMRES return at once, not converged (src/solvers.jl:114-121)."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (3, 3, 3))
    xb = x.copy()
    xb[5] = np.nan
    mf.prepare(xb)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=2)).update()
    assert not pmg.ok
    res = MatrixFreeSolver("fgmres", rtol=1e-8, maxiter=50, preconditioner=pmg).solve(np.zeros(fs.nfree), mf,
                                                                                     np.ones(fs.nfree))
    assert not res.converged and res.iterations == 0


@pytest.mark.parametrize("smoother_storage", ["sym32", "full32", "sym64"])
def test_pmultigrid_mixed_precision(hip_ctx, oracle, smoother_storage):
    """FP32 smoothers and level applies inside an FP64 flexible CG: the Krylov tolerance fixes the answer."""
    from hfk.synthetic import smooth_field

    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-3, smoother_storage=smoother_storage).update()
    assert pmg.ok
    b = np.random.default_rng(12).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=100, preconditioner=pmg).solve(xs, mf, b)
    print(f"mixed {smoother_storage}:", res, flush=True)
    assert res.converged and res.iterations <= 30
    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    xd = spla.spsolve(orc.csc(x).tocsc(), b)
    assert rel(xs, xd) < 1e-7


def test_pmultigrid_h_independent(hip_ctx):
    its = []
    for n in (2, 4, 8):
        fs, fop, cop, mf, x = make_levels(hip_ctx, (n, n, n))
        pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-4).update()
        b = np.random.default_rng(n).standard_normal(fs.nfree)
        res = MatrixFreeSolver("fcg", rtol=1e-8, maxiter=100, preconditioner=pmg).solve(np.zeros(fs.nfree), mf, b)
        assert res.converged
        its.append(res.iterations)
    print("iterations n=2,4,8:", its, flush=True)
    assert max(its) <= 30 and its[2] <= its[1] + 6


def test_nan_state_reports_failure(hip_ctx):
    fs, fop, cop, mf, x = make_levels(hip_ctx, (2, 2, 2))
    xb = x.copy()
    xb[5] = np.nan
    mf.prepare(xb)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=2)).update()
    assert not pmg.ok and not pmg.indefinite
    res = MatrixFreeSolver("fcg", rtol=1e-8, maxiter=50, preconditioner=pmg).solve(np.zeros(fs.nfree), mf, np.ones(fs.nfree))
    assert not res.converged and res.iterations == 0
    z = pmg.ldiv(np.zeros(fs.nfree), np.ones(fs.nfree))
    assert np.isnan(z).all()


def test_indefinite_tangent_reported(hip_ctx):
    """A large smooth state (|grad u| ~ 1) makes the tangent indefinite: the smoother finds a negative Ritz value,
    the p-multigrid is not ok and reports indefinite, the solve returns at once (src/smoothers.jl:48,92-107,
    src/solvers.jl:114-121)."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4), amp=0.05)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4)).update()
    assert not pmg.ok and pmg.indefinite
    res = MatrixFreeSolver("fcg", rtol=1e-8, maxiter=50, preconditioner=pmg).solve(np.zeros(fs.nfree), mf, np.ones(fs.nfree))
    assert not res.converged and res.indefinite and res.iterations == 0


def test_vcycle_symmetric_positive_definite(hip_ctx):
    """The V-cycle is a symmetric positive definite operator on an SPD tangent (test/solvers/test_pmultigrid.jl:74-84):
    Q2 -> Q1 on 2^3 cells at the zero state, coarse level solved to 1e-14 (a fixed operator up to that tolerance);
    the matrix B of the cycle, column by column: ||B - B'|| <= 1e-10 ||B|| and the smallest eigenvalue of its
    symmetric part is positive."""
    fs, fop, cop, mf, x = make_levels(hip_ctx, (2, 2, 2), amp=0.0)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=2), coarse_rtol=1e-14, coarse_maxiter=5000).update()
    assert pmg.ok
    n = fs.nfree
    B = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        B[:, j] = pmg.ldiv(np.zeros(n), e)
    asym = np.linalg.norm(B - B.T) / np.linalg.norm(B)
    lmin = np.linalg.eigvalsh(0.5 * (B + B.T)).min()
    print(f"V-cycle: ||B - B'||/||B|| = {asym:.2e}, min eig {lmin:.3e}", flush=True)
    assert asym <= 1e-10 and lmin > 0


def q3_levels(hip_ctx, n, orders):
    """Levels of orders (3, 1) or (3, 2, 1) on n^3 cells clamped on x = 0, zero state; every coarser level keeps the
    fine 4-point rule (src/pmultigrid.jl:248-251); the last one assembled."""
    ops = []
    for i, o in enumerate(orders):
        sp = hex_space((n, n, n), o, dirichlet=dirichlet_closed_x0)
        kind = "assembled" if i == len(orders) - 1 else "matrixfree"
        ops.append((sp, KernelOperator(HyperelasticForm("neohookean", degree=6), sp, jacobian=kind, ctx=hip_ctx)))
    return ops


@pytest.mark.parametrize("orders", [(3, 1), (3, 2, 1)])
def test_q3_hierarchies(hip_ctx, oracle, orders):
    """Q3 -> Q1 and Q3 -> Q2 -> Q1 (test/solvers/test_pmultigrid.jl:53-66): GMRES(rtol 1e-10) with one V-cycle per
    iteration converges on n = 2 and 4 to the direct solution (1e-7), within 40 iterations, nearly h-independent
    (its(4) <= 1.5 its(2) + 3); the coarse level solved to 1e-13 so that the cycle is a fixed operator."""
    import scipy.sparse.linalg as spla

    its = []
    for n in (2, 4):
        lv = q3_levels(hip_ctx, n, orders)
        fs, fop = lv[0]
        mf = MatrixFreeJacobian(fop).prepare(np.zeros(fs.nfree))
        pmg = PMultigrid([op for _, op in lv], Chebyshev(degree=4), coarse_rtol=1e-13, coarse_maxiter=5000).update()
        assert pmg.ok
        b = np.random.default_rng(n + 30).standard_normal(fs.nfree)
        xs = np.zeros(fs.nfree)
        res = MatrixFreeSolver("gmres", rtol=1e-10, maxiter=200, preconditioner=pmg).solve(xs, mf, b)
        A = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, points1d=4).csc(np.zeros(fs.nfree))
        xd = spla.spsolve(A.tocsc(), b)
        print(f"{orders} n={n}: {res}", flush=True)
        assert res.converged and rel(xs, xd) < 1e-7
        its.append(res.iterations)
    assert max(its) <= 40 and its[1] <= 1.5 * its[0] + 3


def test_plain_cg_with_symmetric_cycle(hip_ctx, oracle):
    """CG (not flexible) with the symmetric cycle (test/solvers/test_pmultigrid.jl:67-71): Q2 -> Q1 on 4^3 cells,
    the coarse level solved to 1e-13: converged, the direct solution to 1e-8."""
    import scipy.sparse.linalg as spla

    fs, fop, cop, mf, x = make_levels(hip_ctx, (4, 4, 4))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-13, coarse_maxiter=5000).update()
    b = np.random.default_rng(41).standard_normal(fs.nfree)
    xs = np.zeros(fs.nfree)
    res = MatrixFreeSolver("cg", rtol=1e-10, maxiter=200, preconditioner=pmg).solve(xs, mf, b)
    from hfk.synthetic import smooth_field

    orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
    print(f"cg + symmetric cycle: {res}", flush=True)
    assert res.converged and not res.indefinite
    assert rel(xs, spla.spsolve(orc.csc(x).tocsc(), b)) < 1e-8


def h_levels(hip_ctx, n, amp=0.01, depth=2, coarse_storage=None):
    """Q2 on n^3 -> Q1 on n^3 (fine 27-point rule) -> Q1 on n^3/8 -> ... (depth halvings, each rediscretised with its
    own 8-point rule and declared a refinement of the next), the last assembled for the direct solve."""
    from hfk.mesh import h_hierarchy
    from hfk.synthetic import smooth_field

    fs = hex_space((n, n, n), 2, dirichlet=dirichlet_closed_x0)
    u = smooth_field(2, amp=amp)
    x = fs.free_values(u)
    fop = KernelOperator(HyperelasticForm("neohookean", degree=4), fs, jacobian="matrixfree",
                         dirichlet_values=fs.dirichlet_values(u), ctx=hip_ctx)
    hs = h_hierarchy(fs, depth, dirichlet=dirichlet_closed_x0)
    ops = [fop]
    for i, sp in enumerate(hs):
        deg = 4 if i == 0 else 2
        kind = "assembled" if i == len(hs) - 1 else "matrixfree"
        ops.append(KernelOperator(HyperelasticForm("neohookean", degree=deg), sp, jacobian=kind, ctx=hip_ctx))
    for a, b in zip(ops[1:-1], ops[2:]):
        a.set_refinement(b)
    mf = MatrixFreeJacobian(fop).prepare(x)
    return fs, ops, mf, x


def test_nested_transfers(hip_ctx):
    """Q1 on 4 x 4 x 4 cells refining Q1 on 2 x 2 x 2 (closed x = 0 Dirichlet face on both): prolongation reproduces
    a trilinear field exactly at the fine free nodes, restriction is its transpose (adjointness), injection of a
    prolongated field gives it back, and levels with other Dirichlet sets or without the declaration are refused."""
    from hfk._lib import HFKError
    from hfk.mesh import dirichlet_closed_x_faces

    fs = hex_space((4, 4, 4), 1, dirichlet=dirichlet_closed_x0)
    cs = hex_space((2, 2, 2), 1, dirichlet=dirichlet_closed_x0)
    form = HyperelasticForm("svk", degree=2)
    fop = KernelOperator(form, fs, jacobian="matrixfree", ctx=hip_ctx)
    cop = KernelOperator(form, cs, jacobian="matrixfree", ctx=hip_ctx)
    with pytest.raises(HFKError):
        Prolongation(cop, fop)  # not declared a refinement
    fop.set_refinement(cop)
    P = Prolongation(cop, fop)
    # trilinear fields that vanish on the Dirichlet face x = 0 (corrections carry no Dirichlet values)
    f = lambda X: X[:, 0:1] * np.stack([1 + 2 * X[:, 1] - 0.5 * X[:, 1] * X[:, 2], X[:, 2] - X[:, 1], 0.3 + X[:, 1] * X[:, 2]],
                                       axis=1)
    assert rel(P @ cs.free_values(f), fs.free_values(f)) < 1e-14
    rng = np.random.default_rng(4)
    xc, yf = rng.standard_normal(cs.nfree), rng.standard_normal(fs.nfree)
    yc = P.restrict(np.zeros(cs.nfree), yf)
    assert abs(yc @ xc - yf @ (P @ xc)) < 1e-13 * np.linalg.norm(yc) * np.linalg.norm(xc)
    xi, di = P.inject(P @ xc, np.zeros(fs.ndir))
    assert rel(xi, xc) < 1e-15
    fs2 = hex_space((4, 4, 4), 1, dirichlet=dirichlet_closed_x_faces)
    f2 = KernelOperator(form, fs2, jacobian="matrixfree", ctx=hip_ctx)
    f2.set_refinement(cop)
    with pytest.raises(HFKError, match="Dirichlet"):
        Prolongation(cop, f2)


@pytest.mark.parametrize("krylov", ["cg", "gmres", "fcg"])
def test_h_hierarchy_solution_and_h_independence(hip_ctx, oracle, krylov):
    """p-multigrid with an h-hierarchy below the Q1 level and a direct solve at the bottom (a fixed, symmetric
    cycle): CG, GMRES and flexible CG reach the direct solution of the Q2 system (1e-8); iterations at most 30 and
    nearly independent of h over n = 4, 8, 16."""
    its = []
    for n in (4, 8, 16):
        fs, ops, mf, x = h_levels(hip_ctx, n, depth={4: 1, 8: 2, 16: 3}[n])
        pmg = PMultigrid(ops, Chebyshev(degree=4), coarse="direct").update()
        assert pmg.ok
        b = np.random.default_rng(n + 50).standard_normal(fs.nfree)
        xs = np.zeros(fs.nfree)
        res = MatrixFreeSolver(krylov, rtol=1e-10, maxiter=200, preconditioner=pmg).solve(xs, mf, b)
        print(f"h-hierarchy {krylov} n={n} levels={len(ops)}: {res}", flush=True)
        assert res.converged and not res.indefinite
        its.append(res.iterations)
        if n == 8:
            from hfk.synthetic import smooth_field

            orc = oracle.OracleOp(fs, "neohookean", 2.0, 1.0, dvals=fs.dirichlet_values(smooth_field(2, amp=0.01)))
            assert rel(xs, spla.spsolve(orc.csc(x).tocsc(), b)) < 1e-8
    assert max(its) <= 30 and its[2] <= its[1] + 6


def test_h_hierarchy_cycle_is_spd(hip_ctx):
    """With the direct coarse solve the V-cycle through the h-levels is a fixed symmetric positive definite
    operator (test/solvers/test_pmultigrid.jl:74-84): Q2 on 4^3, Q1 on 4^3 and 2^3."""
    fs, ops, mf, x = h_levels(hip_ctx, 4, amp=0.0, depth=1)
    pmg = PMultigrid(ops, Chebyshev(degree=2), coarse="direct").update()
    n = fs.nfree
    B = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        B[:, j] = pmg.ldiv(np.zeros(n), e)
    assert np.linalg.norm(B - B.T) <= 1e-12 * np.linalg.norm(B)
    assert np.linalg.eigvalsh(0.5 * (B + B.T)).min() > 0
