"""GPU parity of pointwise laws compiled at run time (hiprtc) -- the reference's generated pointwise integrand
(src/codegen.jl:166-198) evaluated inside the library's staging and contraction kernels.

* The built-in laws written as source (SVK, neo-Hookean, linear) reproduce the built-in kernels and the oracle:
  residual, stored tangents, Jv in every storage, assembled Jacobian values -- FP64 within 1e-12 relative (the
  reference's TOL, test/runtests.jl:9), FP32 storages within 1e-6 of the built-in FP32 storage.
* A law the library does not build in (neo-Hooke with a quadratic volumetric term) is pinned by what the
  mathematics fixes: its tangent at F = I is the linear law's (Jv at the reference state equals the built-in
  linear law's), a rigid rotation is stress free (residual at u = (R - I) X vanishes), Jv equals the central
  difference of the residual (tangent consistent with the stress), and the assembled Jacobian times v equals Jv.
* Compiled code objects are cached by source hash; a law that does not compile raises with the compiler log.
"""
import numpy as np
import pytest

from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk.laws import SOURCES, CompiledLaw
from hfk.mesh import hex_space

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def problem(order, n, perturb, seed, amp=0.02):
    sp = hex_space(n, order, perturb=perturb, seed=seed)
    rng = np.random.default_rng(seed + 100)
    amp = amp / order
    return sp, amp * rng.standard_normal(sp.ndir), amp * rng.standard_normal(sp.nfree), rng


CASES = [
    (1, 2, "svk", (3, 3, 2), 0.0),
    (2, 3, "neohookean", (3, 2, 2), 0.1),
    (2, 3, "linear", (2, 2, 2), 0.1),
    (2, 4, "svk", (2, 2, 2), 0.1),
    (3, 4, "neohookean", (2, 2, 1), 0.1),
]


@pytest.mark.parametrize("order,P,law,n,perturb", CASES)
def test_source_law_matches_builtin_and_oracle(oracle, hip_ctx, order, P, law, n, perturb):
    print(f"compiled law {law} order={order} P={P} n={n}", flush=True)
    sp, dv, x, rng = problem(order, n, perturb, seed=3)
    form = HyperelasticForm(law, 2.0, 1.0, degree=2 * (P - 1))
    ref = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    op.set_law(CompiledLaw.builtin(law, ctx=hip_ctx))
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, points1d=P, dvals=dv)
    b = op.residual(np.zeros(sp.nfree), x)
    assert rel(b, ref.residual(np.zeros(sp.nfree), x)) < TOL
    assert rel(b, orc.residual(x)) < TOL
    dx = rng.standard_normal(sp.nfree)
    for storage in ("full64", "sym64", "full32", "sym32"):
        mf = MatrixFreeJacobian(op, storage).prepare(x)
        mr = MatrixFreeJacobian(ref, storage).prepare(x)
        tol = TOL if storage.endswith("64") else 1e-6
        assert rel(mf @ dx, mr @ dx) < tol, storage
        if storage == "full64":
            prep = orc.prepare(x)
            assert rel(mf.tangents(), oracle.reference_coordinate_tangents(*prep)) < TOL
            assert rel(mf @ dx, orc.jv(prep, dx)) < TOL
    A = op.jacobian(op.allocate_jacobian(), x)
    R = ref.jacobian(ref.allocate_jacobian(), x)
    assert rel(A.data, R.data) < TOL


def test_source_law_with_mu_field(hip_ctx):
    sp, dv, x, rng = problem(2, (3, 2, 2), 0.1, seed=5)
    form = HyperelasticForm("neohookean", 2.0, 1.0, degree=4)
    mu = np.exp(0.5 * rng.standard_normal((sp.ncells, 27)))
    ref = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    ref.set_mu_field(mu)
    op.set_mu_field(mu)
    op.set_law(CompiledLaw.builtin("neohookean", ctx=hip_ctx))
    assert rel(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x)) < TOL
    dx = rng.standard_normal(sp.nfree)
    y = MatrixFreeJacobian(op, "full64").prepare(x) @ dx
    assert rel(y, MatrixFreeJacobian(ref, "full64").prepare(x) @ dx) < TOL
    # back to the built-in law
    op.set_law(None)
    assert rel(MatrixFreeJacobian(op, "sym64").prepare(x) @ dx, y) < TOL


def rotation(axis, angle):
    a = np.asarray(axis, dtype=float)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


@pytest.mark.parametrize("order,P", [(1, 2), (2, 3), (3, 4)])
def test_extra_law_pinned_by_invariants(hip_ctx, order, P):
    print(f"extra law order={order} P={P}", flush=True)
    src, name = SOURCES["neohookean_quadvol"]
    law = CompiledLaw(src, name, ctx=hip_ctx)
    sp, _, _, rng = problem(order, (3, 2, 2), 0.1, seed=11)
    form = HyperelasticForm("svk", 3.0, 0.7, degree=2 * (P - 1))  # built-in law replaced below
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=np.zeros(sp.ndir), ctx=hip_ctx)
    op.set_law(law)
    lin = KernelOperator(HyperelasticForm("linear", 3.0, 0.7, degree=2 * (P - 1)), sp, jacobian="matrixfree",
                         dirichlet_values=np.zeros(sp.ndir), ctx=hip_ctx)
    z = np.zeros(sp.nfree)
    v = rng.standard_normal(sp.nfree)
    # reference state: stress free, tangent = the linear law's
    assert np.abs(op.residual(np.zeros(sp.nfree), z)).max() == 0.0
    y0 = MatrixFreeJacobian(op, "full64").prepare(z) @ v
    assert rel(y0, MatrixFreeJacobian(lin, "full64").prepare(z) @ v) < TOL
    # rigid rotation of 40 degrees: F = R at every point, P = 0
    R = rotation([0.3, -0.5, 0.8], np.deg2rad(40.0))
    rot = lambda X: X @ (R - np.eye(3)).T
    op.set_dirichlet(sp.dirichlet_values(rot))
    xr = sp.free_values(rot)
    br = op.residual(np.zeros(sp.nfree), xr)
    stretch = lambda X: 0.1 * X * X[:, [1, 2, 0]]
    op.set_dirichlet(sp.dirichlet_values(stretch))
    xs = sp.free_values(stretch)
    bs = op.residual(np.zeros(sp.nfree), xs)
    assert np.linalg.norm(br) < 1e-12 * np.linalg.norm(bs)
    # tangent consistent with the stress: central difference of the residual at a stretched state
    h = 1e-6
    y = MatrixFreeJacobian(op, "full64").prepare(xs) @ v
    fd = (op.residual(np.zeros(sp.nfree), xs + h * v) - op.residual(np.zeros(sp.nfree), xs - h * v)) / (2 * h)
    assert rel(fd, y) < 1e-7
    assert rel(MatrixFreeJacobian(op, "sym64").prepare(xs) @ v, y) < TOL
    A = op.jacobian(op.allocate_jacobian(), xs)
    assert rel(A @ v, y) < TOL


def test_cache_and_compile_error(hip_ctx):
    src, name = SOURCES["linear"]
    a = CompiledLaw(src, name, ctx=hip_ctx)
    b = CompiledLaw(src, name, ctx=hip_ctx)
    assert b.from_cache and a.key == b.key
    c = CompiledLaw(src + "\n// another text\n", name, ctx=hip_ctx)
    assert c.key != a.key and not c.from_cache
    with pytest.raises(ValueError, match="no member named"):
        CompiledLaw("struct Empty {};", "Empty", ctx=hip_ctx)


# ---- laws generated from an expression tree (hfk.codegen, the analogue of src/codegen.jl:122-198) ------------

@pytest.mark.parametrize("law,energy", [("svk", "svk_energy"), ("neohookean", "neohooke_energy")])
def test_generated_law_matches_builtin(hip_ctx, law, energy):
    from hfk import codegen as cg

    print(f"generated law {law}", flush=True)
    sp, dv, x, rng = problem(2, (3, 2, 2), 0.1, seed=21)
    form = HyperelasticForm(law, 2.0, 1.0, degree=4)
    ref = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    name = "Gen" + law.capitalize()
    op.set_law(CompiledLaw(cg.from_energy(getattr(cg, energy)).source(name), name, ctx=hip_ctx))
    assert rel(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x)) < TOL
    dx = rng.standard_normal(sp.nfree)
    for storage in ("full64", "sym64"):
        y = MatrixFreeJacobian(op, storage).prepare(x) @ dx
        assert rel(y, MatrixFreeJacobian(ref, storage).prepare(x) @ dx) < TOL, storage


def test_generated_extra_law_invariants(hip_ctx):
    """Compressible Mooney-Rivlin from its energy W = mu/2 (I1-3) + mu/4 (I2-3) - 2 mu lnJ + lam/2 lnJ^2: stress free
    at F = I with the linear law's tangent there (linearising: shear modulus 3 mu / 2, first Lame parameter
    lam + mu), Jv equal to the central difference of the residual, zero residual under a rigid rotation."""
    from hfk import codegen as cg

    def mooney(F, lam, mu):
        C = F.T() @ F
        I1 = cg.tr(C)
        I2 = 0.5 * (I1 * I1 - cg.tr(C @ C))
        lnJ = cg.log(cg.det(F))
        return mu * 0.5 * (I1 - 3.0) + mu * 0.25 * (I2 - 3.0) - 2.0 * mu * lnJ + lam * 0.5 * lnJ * lnJ

    law = CompiledLaw(cg.from_energy(mooney).source("GenMooney"), "GenMooney", ctx=hip_ctx)
    sp, _, _, rng = problem(2, (3, 2, 2), 0.1, seed=23)
    op = KernelOperator(HyperelasticForm("svk", 2.0, 0.8, degree=4), sp, jacobian="matrixfree",
                        dirichlet_values=np.zeros(sp.ndir), ctx=hip_ctx)
    op.set_law(law)
    lin = KernelOperator(HyperelasticForm("linear", 2.0 + 0.8, 1.5 * 0.8, degree=4), sp, jacobian="matrixfree",
                         dirichlet_values=np.zeros(sp.ndir), ctx=hip_ctx)
    z = np.zeros(sp.nfree)
    v = rng.standard_normal(sp.nfree)
    assert np.abs(op.residual(np.zeros(sp.nfree), z)).max() < 1e-14
    assert rel(MatrixFreeJacobian(op, "full64").prepare(z) @ v, MatrixFreeJacobian(lin, "full64").prepare(z) @ v) < TOL
    R = rotation([0.2, 0.9, -0.4], np.deg2rad(35.0))
    rot = lambda X: X @ (R - np.eye(3)).T
    op.set_dirichlet(sp.dirichlet_values(rot))
    br = op.residual(np.zeros(sp.nfree), sp.free_values(rot))
    stretch = lambda X: 0.1 * X * X[:, [1, 2, 0]]
    op.set_dirichlet(sp.dirichlet_values(stretch))
    xs = sp.free_values(stretch)
    bs = op.residual(np.zeros(sp.nfree), xs)
    assert np.linalg.norm(br) < 1e-12 * np.linalg.norm(bs)
    h = 1e-6
    y = MatrixFreeJacobian(op, "full64").prepare(xs) @ v
    fd = (op.residual(np.zeros(sp.nfree), xs + h * v) - op.residual(np.zeros(sp.nfree), xs - h * v)) / (2 * h)
    assert rel(fd, y) < 1e-7


# ---- functions of x materialised on the device (src/coordfunctions.jl:150-213) ------------------------------

def test_mu_function_of_x(oracle, hip_ctx):
    """A shear modulus given as a function of x (HIP source, compiled with hiprtc) is evaluated at the physical
    points of the quadrature points, on non-affine cells: the lognormal random-Fourier-feature field and a
    polynomial without parameters equal the same functions at the oracle's physical points (1e-13), and the
    tangents prepared with the device field equal those with the same field given point by point (1e-12)."""
    from hfk.synthetic import LOGNORMAL_MU_SOURCE, lognormal_mu_params

    sp, dv, x, rng = problem(2, (3, 2, 2), 0.1, seed=9)
    form = HyperelasticForm("neohookean", 2.0, 1.0, degree=4)
    op = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    ref = KernelOperator(form, sp, jacobian="matrixfree", dirichlet_values=dv, ctx=hip_ctx)
    assert np.array_equal(op.mu_field(), np.full((sp.ncells, 27), 1.0))
    X = oracle.physical_points(sp.coords, 3)
    p = lognormal_mu_params(4, sigma=1.0, ell=0.3, features=64, mu0=1.5)
    op.set_mu_function(LOGNORMAL_MU_SOURCE, "lognormal_mu", p)
    w, b = p[3:3 + 192].reshape(64, 3), p[3 + 192:]
    expect = 1.5 * np.exp(np.sqrt(2.0 / 64) * np.cos(X @ w.T + b).sum(axis=2))
    got = op.mu_field()
    assert np.abs(got / expect - 1.0).max() < 1e-13
    ref.set_mu_field(expect)
    dx = rng.standard_normal(sp.nfree)
    y = MatrixFreeJacobian(op, "full64").prepare(x) @ dx
    assert rel(y, MatrixFreeJacobian(ref, "full64").prepare(x) @ dx) < TOL
    assert rel(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x)) < TOL
    poly = "__device__ double poly(const double* x, const double*) { return 1.0 + x[0] * x[0] + 2.0 * x[1] * x[2]; }"
    op.set_mu_function(poly, "poly")
    assert np.abs(op.mu_field() - (1.0 + X[..., 0] ** 2 + 2.0 * X[..., 1] * X[..., 2])).max() < 1e-14
    with pytest.raises(RuntimeError, match="does not compile"):
        op.set_mu_function("__device__ double bad(const double* x, const double* p) { return y; }", "bad")


# ---- update with the same structure: numbers rebound, nothing rebuilt (src/assembly.jl:693-744) ----------------

def test_update_fast_path(hip_ctx, oracle):
    """update(op, res, jac) with other lam, mu: the same operator object with its pattern, the compiled law's module
    kept, and residual, tangents (Jv) and assembled values bitwise equal to a fresh operator of the new numbers and
    equal to the oracle of lam = 3, mu = 0.5 (1e-12); another law builds a new operator."""
    sp, dv, x, rng = problem(2, (3, 2, 2), 0.1, seed=21)
    f1, f2 = HyperelasticForm("neohookean", 2.0, 1.0, degree=4), HyperelasticForm("neohookean", 3.0, 0.5, degree=4)
    law = CompiledLaw.builtin("neohookean", ctx=hip_ctx)
    dx = rng.standard_normal(sp.nfree)
    for compiled in (False, True):
        op = KernelOperator(f1, sp, jacobian="assembled", dirichlet_values=dv, ctx=hip_ctx)
        ref = KernelOperator(f2, sp, jacobian="assembled", dirichlet_values=dv, ctx=hip_ctx)
        if compiled:
            op.set_law(law)
            ref.set_law(law)
        A1 = op.jacobian(op.allocate_jacobian(), x)
        pat = op.pattern()
        assert op.update(f2) is op and op.pattern() is pat
        assert np.array_equal(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x))
        A2 = op.jacobian(op.allocate_jacobian(), x)
        assert np.array_equal(A2.data, ref.jacobian(ref.allocate_jacobian(), x).data)
        assert not np.array_equal(A2.data, A1.data)
        y = MatrixFreeJacobian(op, "full64").prepare(x) @ dx
        assert np.array_equal(y, MatrixFreeJacobian(ref, "full64").prepare(x) @ dx)
        orc = oracle.OracleOp(sp, "neohookean", 3.0, 0.5, dvals=dv)
        rl = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rl(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < 1e-12
        assert rl(y, orc.jv(orc.prepare(x), dx)) < 1e-12
        _, rowval, nzpos = orc.pattern()
        assert rl(A2.data, orc.jacobian(x, nzpos, rowval.size)) < 1e-12
    op2 = op.update(HyperelasticForm("svk", 3.0, 0.5, degree=4))
    assert op2 is not op and op2.form.law == "svk"
