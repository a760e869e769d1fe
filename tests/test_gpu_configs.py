"""The configured workloads at their configured sizes (the oracle is too slow for some of them, so the largest is
checked through size-independent properties).

* The quick start (20 x 10 x 10 Q2 St Venant-Kirchhoff, faces x = 0 clamped and x = 1 moved by (0.05, 0, 0), the
  reference's README.md:61-83): residual and assembled CSC values at x = 0.01 N(0, 1) equal the oracle's (1e-12),
  pattern and nzpos bit for bit.
* The Oseen-BDF2 step on 40^3 x 6 tetrahedra (P2-P1, SUPG/PSPG/LSIC, Laplacian and outlet terms): 31 colours, each a
  valid colouring (no free dof shared by two cells of one colour), nnz 149,854,147, and the residual affine in the
  state with the assembled Jacobian as its slope, r(x) - r(0) = J x (1e-12).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_quickstart_full_size(hip_ctx, oracle):
    from hfk import HyperelasticForm, KernelOperator
    from hfk.mesh import dirichlet_open_x_faces, hex_space
    from hfk.synthetic import normal

    print("quick start 20x10x10 Q2 SVK", flush=True)
    sp = hex_space((20, 10, 10), 2, dirichlet=dirichlet_open_x_faces)
    dv = sp.dirichlet_values(lambda X: np.stack([0.05 * X[:, 0], 0.0 * X[:, 0], 0.0 * X[:, 0]], axis=1))
    assert sp.nfree == 52077
    op = KernelOperator(HyperelasticForm("svk", 2.0, 1.0, degree=4), sp, jacobian="assembled", dirichlet_values=dv,
                        ctx=hip_ctx)
    orc = oracle.OracleOp(sp, "svk", 2.0, 1.0, dvals=dv)
    x = 0.01 * normal(1, sp.nfree)
    assert rel(op.residual(np.zeros(sp.nfree), x), orc.residual(x)) < 1e-12
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = orc.pattern()
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    assert rowval.size == 8990127
    A = op.jacobian(op.allocate_jacobian(), x)
    assert rel(A.data, orc.jacobian(x, p2, r2.size)) < 1e-12


def test_oseen_40cube_properties(hip_ctx):
    from hfk.oseen import OseenOperator
    from hfk.tetmesh import divfree_field, outlet_faces, p2_laplacian, tet_p2p1_space

    print("oseen 40^3 x 6 tetrahedra", flush=True)
    sp = tet_p2p1_space((40, 40, 40))
    wraw = divfree_field(5)
    scale = 1.0 / np.sqrt((wraw(sp.node_xyz) ** 2).sum(axis=1).mean())
    w = lambda X: scale * wraw(X)
    gin = lambda X: np.stack([16 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2]), 0 * X[:, 0], 0 * X[:, 0]], 1)
    xd = sp.node_xyz[sp.dir_node_of]
    dv = np.where((xd[:, 0] == 0.0)[:, None], gin(xd), 0.0).reshape(-1)
    op = OseenOperator(sp, rho=1.0, mu=0.01, dt=0.01, bdf2=True, dirichlet_values=dv, ctx=hip_ctx)
    op.set_steady_advection(w)
    un = np.zeros(sp.nfree)
    un[:sp.nfree_u] = w(sp.node_xyz[sp.free_node_of]).reshape(-1)
    op.set_laplacian(p2_laplacian(sp, un, sp.dirichlet_values(w)))
    op.set_outlet(outlet_faces(sp), w, Pc=0.7, ZA=2.3)
    colptr, rowval, nzpos = op.pattern()
    del nzpos
    assert rowval.size == 149854147
    colors = op.colors()
    assert colors.max() == 31 and colors.min() == 1
    ids = sp.cell_dofs
    for c in range(1, 32):
        f = ids[colors == c]
        f = f[f > 0]
        assert np.unique(f).size == f.size, c  # the cells of one colour share no free dof
    A = op.jacobian(op.allocate_jacobian())
    x = np.random.default_rng(5).standard_normal(sp.nfree)
    r0 = op.residual(np.zeros(sp.nfree), np.zeros(sp.nfree))
    r1 = op.residual(np.zeros(sp.nfree), x)
    Ax = A @ x
    print(f"oseen 40: |r(x) - r(0) - A x| / |A x| = {rel(r1 - r0, Ax):.2e}", flush=True)
    assert rel(r1 - r0, Ax) < 1e-12
