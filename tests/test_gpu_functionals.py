"""GPU parity of the scalar functionals (KernelFunctional, src/functionals.jl:9-154) against the oracle and the
reference's own functional tests (test/solvers/test_functionals.jl).

* Energy (reduce = sum): equal to the oracle's total energy (1e-12) for the three built-in laws on Q1-Q3 non-affine
  cells with Dirichlet values, and with a shear-modulus field; its directional derivative equals residual . dx
  (central differences, 1e-6, test_functionals.jl:33-37); bitwise reproducible.  Compiled laws (sources and a law
  generated from its energy) give the built-in law's energy; a compiled law without energy() is refused.
* Extremum (reduce = min / max) of g(Gu, Gdu) = det(I + Gu) / (p0 + |tr Gdu| + p1 Gdu:Gdu) (the function of
  test_functionals.jl:59-78, with its numbers as parameters) over the quadrature points: equal to the same function
  of the oracle's point gradients (1e-13), with the increment changing between calls and with zero or the
  operator's Dirichlet values for the increment.
"""
import numpy as np
import pytest

from hfk import HyperelasticForm, KernelOperator
from hfk.functionals import KernelFunctional
from hfk.mesh import dirichlet_closed_x0, hex_space

pytestmark = pytest.mark.gpu


def problem(hip_ctx, oracle, order, law, n, seed, amp=0.02):
    sp = hex_space(n, order, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=seed)
    rng = np.random.default_rng(seed + 50)
    dv = amp / order * rng.standard_normal(sp.ndir)
    x = amp / order * rng.standard_normal(sp.nfree)
    op = KernelOperator(HyperelasticForm(law, 2.0, 1.0, degree=2 * order), sp, jacobian="matrixfree",
                        dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    return sp, dv, x, op, orc, rng


@pytest.mark.parametrize("order,law,n", [(1, "svk", (3, 3, 2)), (2, "neohookean", (3, 2, 2)), (2, "linear", (2, 2, 2)),
                                         (3, "svk", (2, 2, 1))])
def test_energy_matches_oracle_and_residual(hip_ctx, oracle, order, law, n):
    print(f"energy Q{order} {law}", flush=True)
    sp, dv, x, op, orc, rng = problem(hip_ctx, oracle, order, law, n, seed=4)
    fun = KernelFunctional(op)
    E = fun(x)
    assert E == pytest.approx(orc.energy(x), rel=1e-12)
    assert fun(x) == E  # fixed summation order
    dx = rng.standard_normal(sp.nfree)
    h = 1e-6
    fd = (fun(x + h * dx) - fun(x - h * dx)) / (2 * h)
    assert fd == pytest.approx(op.residual(np.zeros(sp.nfree), x) @ dx, rel=1e-6)
    mu = np.exp(0.5 * rng.standard_normal((sp.ncells, (order + 1) ** 3)))
    op.set_mu_field(mu)
    orc.set_mu_field(mu)
    assert fun(x) == pytest.approx(orc.energy(x), rel=1e-12)


NO_ENERGY = r"""
struct LinearNoEnergy {
  double G[9], lam, mu, tr;
  __device__ void setup(const double* Fin, double l, double m) {
    lam = l; mu = m;
    for (int c = 0; c < 3; ++c)
      for (int J = 0; J < 3; ++J) G[c + 3 * J] = Fin[c + 3 * J] - (c == J ? 1.0 : 0.0);
    tr = G[0] + G[4] + G[8];
  }
  __device__ double stress(int c, int J) const { return (c == J ? lam * tr : 0.0) + mu * (G[c + 3 * J] + G[J + 3 * c]); }
  __device__ double tangent(int c, int J, int k, int L) const {
    return (c == J && k == L ? lam : 0.0) + (c == k && J == L ? mu : 0.0) + (J == k && c == L ? mu : 0.0);
  }
};
"""


@pytest.mark.parametrize("law", ["svk", "neohookean", "linear"])
def test_energy_of_compiled_laws(hip_ctx, oracle, law):
    from hfk import codegen as cg
    from hfk.laws import SOURCES, CompiledLaw

    sp, dv, x, op, orc, rng = problem(hip_ctx, oracle, 2, law, (3, 2, 2), seed=8)
    fun = KernelFunctional(op)
    E = fun(x)
    src, name = SOURCES[law]
    op.set_law(CompiledLaw(src, name, ctx=hip_ctx))
    assert fun(x) == pytest.approx(E, rel=1e-13)
    if law != "linear":
        energy = cg.svk_energy if law == "svk" else cg.neohooke_energy
        op.set_law(CompiledLaw(cg.from_energy(energy).source("Gen" + name), "Gen" + name, ctx=hip_ctx))
        assert fun(x) == pytest.approx(E, rel=1e-12)
    else:
        op.set_law(CompiledLaw(NO_ENERGY, "LinearNoEnergy", ctx=hip_ctx))
        with pytest.raises(RuntimeError, match="defines no"):
            fun(x)
    op.set_law(None)
    assert fun(x) == E


G_SOURCE = r"""
__device__ double injectivity_probe(const double* Gu, const double* Gdu, const double* p) {
  double F[9], dd = 0;
  for (int i = 0; i < 9; ++i) {
    F[i] = Gu[i] + (i % 4 == 0 ? 1.0 : 0.0);
    dd += Gdu[i] * Gdu[i];
  }
  return hfk_det3(F) / (p[0] + fabs(Gdu[0] + Gdu[4] + Gdu[8]) + p[1] * dd);
}
"""


def g_of(Gu, Gdu, p):
    F = np.eye(3) + Gu
    return np.linalg.det(F) / (p[0] + np.abs(np.trace(Gdu, axis1=-2, axis2=-1)) + p[1] * (Gdu * Gdu).sum(axis=(-2, -1)))


@pytest.mark.parametrize("order", [2, 3])
def test_extremum_matches_oracle(hip_ctx, oracle, order):
    print(f"extremum Q{order}", flush=True)
    n = (3, 2, 2) if order == 2 else (2, 2, 1)
    sp, dv, x, op, orc, rng = problem(hip_ctx, oracle, order, "svk", n, seed=12, amp=0.1)
    p = np.array([1.0, 0.1])
    fmin = KernelFunctional(op, "min", G_SOURCE, "injectivity_probe", params=p)
    fmax = KernelFunctional(op, "max", G_SOURCE, "injectivity_probe", params=p)
    orc0 = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=np.zeros(sp.ndir))
    for k in range(3):
        x = 0.1 / order * rng.standard_normal(sp.nfree)
        dx = rng.standard_normal(sp.nfree)
        # the increment with the operator's Dirichlet values (an FE function on the same space), then with zero ones
        for dd, o in ((dv, orc), (None, orc0)):
            fmin.set_increment(dx, dd)
            fmax.set_increment(dx, dd)
            g = g_of(orc.point_gradients(x), o.point_gradients(dx), p)
            vmin, vmax = fmin(x), fmax(x)
            assert vmin == pytest.approx(g.min(), rel=1e-13) and vmax == pytest.approx(g.max(), rel=1e-13)
            assert vmin < vmax


@pytest.mark.parametrize("order,law,n", [(1, "svk", (3, 3, 2)), (2, "neohookean", (3, 2, 2)), (3, "svk", (2, 2, 1))])
def test_traction_load(hip_ctx, oracle, order, law, n):
    """A dead traction on two boundary faces (a boundary group of the residual and of the energy): the load vector
    equals the oracle's (1e-13); residual! = residual without the load - Λ f on the free rows; the energy = the
    oracle's energy - Λ (f . x + f_D . u_D) (1e-12) and its central difference = residual . dx (1e-6); the assembled
    Jacobian does not change; Λ = 0 and removing the faces give back the volume terms."""
    from hfk.mesh import boundary_faces

    print(f"traction Q{order} {law}", flush=True)
    sp, dv, x, op, orc, rng = problem(hip_ctx, oracle, order, law, n, seed=31)
    fc1, fl1 = boundary_faces(sp, 0, 1)
    fc2, fl2 = boundary_faces(sp, 2, 0)
    fc, fl = np.concatenate([fc1, fc2]), np.concatenate([fl1, fl2])
    tr = np.concatenate([np.tile([0.3, -0.2, 0.7], (fc1.size, 1)), np.tile([-1.1, 0.4, 0.25], (fc2.size, 1))])
    fun = KernelFunctional(op)
    b0 = op.residual(np.zeros(sp.nfree), x)
    A0 = op.jacobian(op.allocate_jacobian(), x).data.copy()
    E0 = fun(x)
    op.set_traction(fc, fl, tr)
    lam = 1.7
    op.set_load_factor(lam)
    ff, fd = orc.face_load(fc, fl, tr)
    gf, gd = op.traction_load()
    scale = np.abs(np.concatenate([ff, fd])).max()
    assert np.abs(gf - ff).max() <= 1e-13 * scale and np.abs(gd - fd).max() <= 1e-13 * scale
    b = op.residual(np.zeros(sp.nfree), x)
    assert np.abs(b - (b0 - lam * gf)).max() <= 1e-14 * np.abs(b).max()
    E = fun(x)
    assert E == pytest.approx(orc.energy(x) - lam * (ff @ x + fd @ dv), rel=1e-12, abs=1e-14)
    dx = rng.standard_normal(sp.nfree)
    h = 1e-6
    assert (fun(x + h * dx) - fun(x - h * dx)) / (2 * h) == pytest.approx(b @ dx, rel=1e-6)
    assert np.array_equal(op.jacobian(op.allocate_jacobian(), x).data, A0)
    op.set_load_factor(0.0)
    assert np.array_equal(op.residual(np.zeros(sp.nfree), x), b0) and fun(x) == E0
    op.set_load_factor(1.0)
    op.set_traction(None, None, None)
    assert np.array_equal(op.residual(np.zeros(sp.nfree), x), b0) and fun(x) == E0


@pytest.mark.parametrize("order,law,n", [(1, "svk", (3, 3, 2)), (2, "neohookean", (3, 2, 2)), (3, "linear", (2, 2, 1))])
def test_robin_term(hip_ctx, oracle, order, law, n):
    """A Robin (spring) term k int u.v dGamma on two boundary faces with different stiffnesses, against the oracle's
    face mass matrix M over all dofs: residual = volume + (M u)_free with the Dirichlet values lifted; assembled
    Jacobian = volume + M_ff; Jv (full64 and sym64 storages) = volume + M_ff dx; the inverted node blocks = the
    inverses of the 3 x 3 node blocks of the oracle's Jacobian + M_ff; energy = volume + u.M u / 2 and its central
    difference = residual . dx; removing the faces gives back the volume terms (all 1e-13 / 1e-12)."""
    from hfk import MatrixFreeJacobian
    from hfk.mesh import boundary_faces
    from hfk.solvers import BlockJacobi

    print(f"robin Q{order} {law}", flush=True)
    sp, dv, x, op, orc, rng = problem(hip_ctx, oracle, order, law, n, seed=37)
    fc1, fl1 = boundary_faces(sp, 0, 1)
    fc2, fl2 = boundary_faces(sp, 2, 0)
    fc, fl = np.concatenate([fc1, fc2]), np.concatenate([fl1, fl2])
    k = np.concatenate([np.full(fc1.size, 1.3), np.full(fc2.size, 0.4)])
    nf = sp.nfree
    M = orc.face_mass(fc, fl, k)
    Mff = M[:nf, :nf]
    u = np.concatenate([x, dv])
    fun = KernelFunctional(op)
    dx = rng.standard_normal(nf)
    b0 = op.residual(np.zeros(nf), x)
    A0 = op.jacobian(op.allocate_jacobian(), x).toarray()
    E0 = fun(x)
    y0 = {st: MatrixFreeJacobian(op, st).prepare(x) @ dx for st in ("full64", "sym64")}
    op.set_robin(fc, fl, k)
    b = op.residual(np.zeros(nf), x)
    assert np.abs(b - (b0 + (M @ u)[:nf])).max() <= 1e-13 * np.abs(b).max()
    A = op.jacobian(op.allocate_jacobian(), x).toarray()
    assert np.abs(A - (A0 + Mff)).max() <= 1e-13 * np.abs(A).max()
    for st in ("full64", "sym64"):
        mf = MatrixFreeJacobian(op, st).prepare(x)
        y = mf @ dx
        assert np.abs(y - (y0[st] + Mff @ dx)).max() <= 1e-13 * np.abs(y).max()
    mf = MatrixFreeJacobian(op, "full64").prepare(x)
    bj = BlockJacobi(mf)
    assert bj.ok
    dofs, Binv = bj.blocks()
    Atot = orc.csc(x).toarray() + Mff
    for blk in range(dofs.shape[0]):
        ids = dofs[blk]
        B = np.eye(3)
        fr = ids > 0
        B[np.ix_(fr, fr)] = Atot[np.ix_(ids[fr] - 1, ids[fr] - 1)]
        Bi = np.linalg.inv(B).reshape(-1, order="F")
        assert np.abs(Binv[blk] - Bi).max() <= 1e-12 * np.abs(Bi).max()
    E = fun(x)
    assert E == pytest.approx(E0 + 0.5 * u @ M @ u, rel=1e-12)
    h = 1e-6
    assert (fun(x + h * dx) - fun(x - h * dx)) / (2 * h) == pytest.approx(b @ dx, rel=1e-6)
    op.set_robin(None, None, None)
    assert np.array_equal(op.residual(np.zeros(nf), x), b0) and fun(x) == E0
