"""Status codes of the C ABI on misuse, with the reference's exception kinds (the Python mirror raises HFKError
naming ArgumentError / DimensionMismatch / ErrorException): BlockJacobi and the p-multigrid before prepare!
(test_matrixfree_solver.jl:66-68, test_pmultigrid.jl:228), out-of-range settings of the Krylov solve, the coarse
network and the Oseen operator.  Every call fails before launching any kernel."""
import ctypes as C

import numpy as np
import pytest

import hfk
from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk._lib import HFKError, check
from hfk.mesh import dirichlet_closed_x0, hex_space
from hfk.solvers import BlockJacobi, Chebyshev, PMultigrid

pytestmark = pytest.mark.gpu


def levels(hip_ctx):
    fs = hex_space((2, 2, 2), 2, dirichlet=dirichlet_closed_x0)
    cs = hex_space((2, 2, 2), 1, dirichlet=dirichlet_closed_x0)
    fop = KernelOperator(HyperelasticForm("svk", degree=4), fs, jacobian="matrixfree", ctx=hip_ctx)
    cop = KernelOperator(HyperelasticForm("svk", degree=4), cs, jacobian="assembled", ctx=hip_ctx)
    return fs, cs, fop, cop


def test_setup_before_prepare(hip_ctx):
    fs, cs, fop, cop = levels(hip_ctx)
    mf = MatrixFreeJacobian(fop)
    with pytest.raises(HFKError, match="ErrorException"):
        BlockJacobi(mf)
    with pytest.raises(HFKError, match="ErrorException"):
        PMultigrid([fop, cop], Chebyshev()).update()
    with pytest.raises(HFKError, match="ErrorException"):
        mf @ np.zeros(fs.nfree)


def test_out_of_range_settings(hip_ctx):
    fs, cs, fop, cop = levels(hip_ctx)
    mf = MatrixFreeJacobian(fop).prepare(np.zeros(fs.nfree))
    lib = fop.lib
    st = np.zeros(3, dtype=np.int32)
    rn = np.zeros(2)
    b = np.ones(fs.nfree)
    x = np.zeros(fs.nfree)
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    for kind, restart in ((4, 50), (2, 0), (-1, 50)):
        rc = lib.hfk_krylov_solve(fop.handle, None, kind, 0, P(b, C.c_double), P(x, C.c_double), 1e-8, 0.0, 10, restart,
                                  P(st, C.c_int32), P(rn, C.c_double))
        assert rc == -1  # ArgumentError
    pmg = PMultigrid([fop, cop], Chebyshev())
    with pytest.raises(HFKError, match="ArgumentError"):
        check(lib.hfk_pmg_set_coarse_network(pmg.handle, None, 7, 0, 1), fop.ctx.handle)
    with pytest.raises(HFKError, match="ArgumentError"):
        check(lib.hfk_pmg_set_coarse_network(pmg.handle, None, 1, -1, 1), fop.ctx.handle)


def test_oseen_and_network_arguments(hip_ctx):
    from hfk.learned import CoarseNetwork
    from hfk.oseen import OseenOperator
    from hfk.tetmesh import tet_p2p1_space, tet_quadrature_deg5

    sp = tet_p2p1_space((1, 1, 1))
    qp, qw = tet_quadrature_deg5()
    qp40 = np.tile(qp, (3, 1))[:40]
    qw40 = np.tile(qw, 3)[:40]
    with pytest.raises(HFKError, match="ArgumentError"):
        OseenOperator(sp, quadrature=(qp40, qw40), ctx=hip_ctx)  # more than 32 points
    with pytest.raises(HFKError, match="ArgumentError"):
        OseenOperator(sp, mu=0.0, ctx=hip_ctx)
    bad = tet_p2p1_space((1, 1, 1))
    bad.cell_dofs = bad.cell_dofs.copy()
    bad.cell_dofs[0, 0] = bad.nfree + 5
    with pytest.raises(HFKError, match="DimensionMismatch"):
        OseenOperator(bad, ctx=hip_ctx)
    op = OseenOperator(sp, ctx=hip_ctx)
    with pytest.raises(HFKError, match="ErrorException"):  # no point data yet
        op.residual(np.zeros(sp.nfree), np.zeros(sp.nfree))
    with pytest.raises(ValueError):
        op.set_dirichlet(np.zeros(sp.ndir + 1))
    cs = hex_space((4, 4, 4), 1, dirichlet=dirichlet_closed_x0)  # 5^3 lattice: too small for three halvings
    with pytest.raises(HFKError, match="ArgumentError"):
        CoarseNetwork(hip_ctx, cs)


def test_duplicate_faces_rejected(hip_ctx):
    """A (cell, local face) pair given twice would put both copies in one launch group, whose faces must share no
    dof (two threads would add into the same entries): the traction and Robin set-ups refuse it."""
    from hfk.mesh import boundary_faces

    fs, cs, fop, cop = levels(hip_ctx)
    fc, fl = boundary_faces(fs, 0, 1)
    dup_c, dup_l = np.concatenate([fc, fc[:1]]), np.concatenate([fl, fl[:1]])
    with pytest.raises(HFKError, match="ArgumentError"):
        fop.set_traction(dup_c, dup_l, [1.0, 0.0, 0.0])
    with pytest.raises(HFKError, match="ArgumentError"):
        fop.set_robin(dup_c, dup_l, 1.0)
    fop.set_traction(fc, fl, [1.0, 0.0, 0.0])  # the same faces once each are fine
    fop.set_robin(fc, fl, 1.0)
