"""GPU parity of the Oseen-BDF2 step on tetrahedra (P2-P1, SUPG/PSPG/LSIC; test/common.jl:165-228) against the
oracle (tests run as test/test_cfd.jl does against Gridap): sparsity and nzpos bit-exact, colours equal to the
greedy first fit, residual and CSC Jacobian values within 1e-12 relative, the residual affine in the state
(r(x) = J x + r(0), the lifted system A = J, b = -r(0)), the fused residual+Jacobian pass equal to the two
separate ones, and the full step with the Laplacian in the stabilisation load and the outlet terms.
"""
import numpy as np
import pytest

from hfk.oseen import OseenOperator
from hfk.tetmesh import divfree_field, tet_p2p1_space, tet_quadrature_deg5

pytestmark = pytest.mark.gpu

TOL = 1e-12


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def setup(hip_ctx, n, bdf2, seed):
    sp = tet_p2p1_space(n, domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    qp, qw = tet_quadrature_deg5()
    rng = np.random.default_rng(seed)
    fields = rng.standard_normal((sp.ncells, len(qw), 9))
    dv = rng.standard_normal(sp.ndir)
    op = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.01, bdf2=bdf2, dirichlet_values=dv, ctx=hip_ctx)
    op.set_fields(fields[..., 0:3], fields[..., 3:6], fields[..., 6:9])
    return sp, qp, qw, fields, dv, op, rng


@pytest.mark.parametrize("n,bdf2", [((2, 2, 1), False), ((3, 2, 2), True)])
def test_oseen_pattern_residual_jacobian(hip_ctx, n, bdf2):
    from oracle import oseen as O

    print(f"oseen n={n} bdf2={bdf2}", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, n, bdf2, seed=7)
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = O.pattern(sp.cell_dofs, sp.nfree)
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    assert np.array_equal(op.colors(), O.greedy_colors(sp.cell_dofs, sp.nfree))
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0)
    A = op.jacobian(op.allocate_jacobian())
    assert rel(A.data, O.jacobian_values(K, p2, r2.size)) < TOL
    x = rng.standard_normal(sp.nfree)
    b = op.residual(np.zeros(sp.nfree), x)
    assert rel(b, O.residual(sp, x, dv, K, load)) < TOL
    b0 = op.residual(np.zeros(sp.nfree), np.zeros(sp.nfree))
    assert rel(b - b0, A @ x) < 1e-11


def test_oseen_fused_and_steady_advection(hip_ctx):
    import torch

    from oracle import oseen as O

    sp = tet_p2p1_space((3, 3, 2))
    op = OseenOperator(sp, rho=1.0, mu=0.01, dt=0.01, bdf2=True, ctx=hip_ctx)
    w = divfree_field(5)
    gin = lambda X: np.stack([16 * X[:, 1] * (1 - X[:, 1]) * X[:, 2] * (1 - X[:, 2]), 0 * X[:, 0], 0 * X[:, 0]], 1)
    xd = sp.node_xyz[sp.dir_node_of]
    dv = np.where((xd[:, 0] == 0.0)[:, None], gin(xd), 0.0).reshape(-1)
    op.set_dirichlet(dv)
    op.set_steady_advection(w)
    # oracle with the same point data
    qp, qw = tet_quadrature_deg5()
    wq = w(sp.quadrature_points(qp).reshape(-1, 3)).reshape(sp.ncells, len(qw), 3)
    fields = np.concatenate([wq, wq, op.c0 * wq], axis=2)
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.01, op.c0)
    colptr, rowval, nzpos = op.pattern()
    x = np.random.default_rng(2).standard_normal(sp.nfree)
    ref_b = O.residual(sp, x, dv, K, load)
    ref_nz = O.jacobian_values(K, nzpos, rowval.size)
    dev = torch.device("cuda:0")
    xt = torch.from_numpy(x).to(dev)
    bt = torch.zeros(sp.nfree, dtype=torch.float64, device=dev)
    nt = torch.zeros(rowval.size, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    import ctypes as C

    from hfk._lib import check

    check(op.lib.hfk_oseen_residual_jacobian_dev(op.handle, C.c_void_p(xt.data_ptr()), C.c_void_p(bt.data_ptr()),
                                                 C.c_void_p(nt.data_ptr()), None), op.ctx.handle)
    torch.cuda.synchronize()
    assert rel(bt.cpu().numpy(), ref_b) < TOL
    assert rel(nt.cpu().numpy(), ref_nz) < TOL
    assert rel(op.residual(np.zeros(sp.nfree), x), ref_b) < TOL


@pytest.mark.parametrize("n,bdf2", [((3, 2, 2), True), ((4, 2, 3), False)])
def test_oseen_full_step_with_outlet_and_laplacian(hip_ctx, n, bdf2):
    """Every term of the reference's step: volume, the Laplacian of un in the stabilisation load, and the outlet
    backflow (w random: inflow and outflow points), Robin and pressure terms; then the pressure alone updated."""
    from oracle import oseen as O

    from hfk.tetmesh import outlet_faces

    print(f"oseen full step n={n} bdf2={bdf2}", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, n, bdf2, seed=11)
    lap = rng.standard_normal((sp.ncells, len(qw), 3))
    op.set_laplacian(lap)
    of = outlet_faces(sp)
    wb = rng.standard_normal((len(of.cell), of.xb.shape[1], 3))
    op.set_outlet(of, wb, Pc=0.7, ZA=2.3)
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0, lap=lap)
    xb = np.concatenate([of.xb, wb], axis=2)
    Kf, Lf = O.outlet_terms(of.xo, xb, of.normal, 1.0, 0.7, 2.3)
    Kt, lt = O.add_faces(K, load, of.cell, Kf, Lf)
    colptr, rowval, nzpos = op.pattern()
    A = op.jacobian(op.allocate_jacobian())
    assert rel(A.data, O.jacobian_values(Kt, nzpos, rowval.size)) < TOL
    x = rng.standard_normal(sp.nfree)
    b = op.residual(np.zeros(sp.nfree), x)
    assert rel(b, O.residual(sp, x, dv, Kt, lt)) < TOL
    assert rel(b - op.residual(np.zeros(sp.nfree), np.zeros(sp.nfree)), A @ x) < 1e-11
    # the Windkessel pressure of the next step
    op.set_outlet_pressure(1.9)
    _, Lf2 = O.outlet_terms(of.xo, xb, of.normal, 1.0, 1.9, 2.3)
    _, lt2 = O.add_faces(K, load, of.cell, Kf, Lf2)
    assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, Kt, lt2)) < TOL
    # terms removed again: the volume step alone
    op.set_outlet(None)
    op.set_laplacian(None)
    K0, l0 = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0)
    assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, K0, l0)) < TOL


def test_oseen_atomic_single_launch(hip_ctx):
    """One launch with FP64 atomic adds (cells in mesh order, a block per XCD) against the colour-by-colour
    assembly: the same residual and CSC values up to the order of the additions."""
    from hfk.tetmesh import outlet_faces

    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, (4, 3, 2), True, seed=13)
    op.set_laplacian(rng.standard_normal((sp.ncells, len(qw), 3)))
    of = outlet_faces(sp)
    op.set_outlet(of, rng.standard_normal((len(of.cell), of.xb.shape[1], 3)), Pc=0.7, ZA=2.3)
    x = rng.standard_normal(sp.nfree)
    b0 = op.residual(np.zeros(sp.nfree), x)
    A0 = op.jacobian(op.allocate_jacobian())
    op.set_deterministic(False)
    b1 = op.residual(np.zeros(sp.nfree), x)
    A1 = op.jacobian(op.allocate_jacobian())
    assert rel(b1, b0) < TOL and rel(A1.data, A0.data) < TOL
    op.set_deterministic(True)
    assert np.array_equal(op.residual(np.zeros(sp.nfree), x), b0)  # colour order: bitwise reproducible


@pytest.mark.parametrize("bdf2", [True, False])
def test_oseen_step_fields_time_loop(hip_ctx, bdf2):
    """The point data of each step made on the device from the FE functions of the two previous steps (w, un, f,
    the Laplacian of un and w at the outlet backflow points), over two steps of the time loop of
    test/test_cfd.jl:47-75 (previous states shifted, Dirichlet values of the previous velocity rescaled, the
    Windkessel pressure changed): residual and CSC values equal to the oracle's system built from the same FE
    functions (1e-12); the Laplacian term removed on request; the device-pointer variant gives the same values."""
    import torch

    from oracle import oseen as O

    from hfk.tetmesh import outlet_faces

    print(f"oseen step fields bdf2={bdf2}", flush=True)
    sp = tet_p2p1_space((3, 2, 2), domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    qp, qw = tet_quadrature_deg5()
    rng = np.random.default_rng(17)
    dv = rng.standard_normal(sp.ndir)
    op = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.01, bdf2=bdf2, dirichlet_values=dv, ctx=hip_ctx)
    of = outlet_faces(sp)
    op.set_outlet(of, None, Pc=0.7, ZA=2.3)
    aw, bw, c1, c2 = op.step_coefficients()
    un, unm1 = rng.standard_normal(sp.nfree), rng.standard_normal(sp.nfree)
    und, unm1d = rng.standard_normal(sp.ndir), rng.standard_normal(sp.ndir)
    x = rng.standard_normal(sp.nfree)
    colptr, rowval, nzpos = op.pattern()

    def oracle_system(Pc, with_lap=True):
        fields, lap = O.step_fields(sp.cell_dofs, sp.coords, qp, un, und, unm1, unm1d, aw, bw, c1, c2)
        K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0, lap=lap if with_lap else None)
        wb = O.face_velocity(sp.cell_dofs, of.cell, of.xb[..., :3], un, und, unm1, unm1d, aw, bw)
        Kf, Lf = O.outlet_terms(of.xo, np.concatenate([of.xb, wb], axis=2), of.normal, 1.0, Pc, 2.3)
        return O.add_faces(K, load, of.cell, Kf, Lf)

    for step, Pc in enumerate((0.7, -1.3)):
        op.set_outlet_pressure(Pc)
        op.step_fields(un, und, unm1, unm1d)
        Kt, lt = oracle_system(Pc)
        b = op.residual(np.zeros(sp.nfree), x)
        assert rel(b, O.residual(sp, x, dv, Kt, lt)) < TOL
        A = op.jacobian(op.allocate_jacobian())
        assert rel(A.data, O.jacobian_values(Kt, nzpos, rowval.size)) < TOL
        # the next step: previous states shifted, new u^n, Dirichlet values of the previous velocities rescaled
        unm1, un = un, rng.standard_normal(sp.nfree)
        unm1d, und = und, 0.4 * und
    # the Laplacian term removed
    op.step_fields(un, und, unm1, unm1d, laplacian=False)
    Kt, lt = oracle_system(-1.3, with_lap=False)
    b0 = op.residual(np.zeros(sp.nfree), x)
    assert rel(b0, O.residual(sp, x, dv, Kt, lt)) < TOL
    # device pointers, asynchronous on the context's stream
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in (un, und, unm1, unm1d)]
    torch.cuda.synchronize()
    op.step_fields_dev(*[v.data_ptr() for v in t])
    Kt, lt = oracle_system(-1.3)
    b1 = op.residual(np.zeros(sp.nfree), x)
    assert rel(b1, O.residual(sp, x, dv, Kt, lt)) < TOL
    op.step_fields(un, und, unm1, unm1d)
    assert np.array_equal(op.residual(np.zeros(sp.nfree), x), b1)


def test_oseen_update_bdf2_switch(hip_ctx):
    """The switch from backward Euler to BDF2 through update (test/test_cfd.jl:66-75): the same operator and pattern,
    residual and CSC values bitwise equal to an operator created with BDF2."""
    print("oseen update bdf2 switch", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, (3, 2, 2), False, seed=19)
    ref = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.01, bdf2=True, dirichlet_values=dv, ctx=hip_ctx)
    ref.set_fields(fields[..., 0:3], fields[..., 3:6], fields[..., 6:9])
    x = rng.standard_normal(sp.nfree)
    pat = op.pattern()
    b1 = op.residual(np.zeros(sp.nfree), x)
    assert op.update(bdf2=True) is op and op.pattern() is pat
    b2 = op.residual(np.zeros(sp.nfree), x)
    assert np.array_equal(b2, ref.residual(np.zeros(sp.nfree), x)) and not np.array_equal(b1, b2)
    assert np.array_equal(op.jacobian(op.allocate_jacobian()).data, ref.jacobian(ref.allocate_jacobian()).data)
    assert op.step_coefficients() == ref.step_coefficients()


def test_oseen_two_faces_of_one_cell(hip_ctx):
    """Faces of one cell go to different launches: with every outlet face listed twice (two faces per cell) the
    residual and CSC values equal the oracle's system with each face term counted twice."""
    import dataclasses

    from oracle import oseen as O

    from hfk.tetmesh import outlet_faces

    print("oseen two faces per cell", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, (3, 2, 2), True, seed=23)
    of = outlet_faces(sp)
    twice = dataclasses.replace(of, cell=np.concatenate([of.cell, of.cell]), normal=np.concatenate([of.normal] * 2),
                                xo=np.concatenate([of.xo] * 2), xb=np.concatenate([of.xb] * 2),
                                xb_phys=np.concatenate([of.xb_phys] * 2))
    wb = rng.standard_normal((len(of.cell), of.xb.shape[1], 3))
    op.set_outlet(twice, np.concatenate([wb, wb]), Pc=0.7, ZA=2.3)
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0)
    Kf, Lf = O.outlet_terms(of.xo, np.concatenate([of.xb, wb], axis=2), of.normal, 1.0, 0.7, 2.3)
    Kt, lt = O.add_faces(K, load, np.concatenate([of.cell, of.cell]), np.concatenate([Kf, Kf]), np.concatenate([Lf, Lf]))
    colptr, rowval, nzpos = op.pattern()
    x = rng.standard_normal(sp.nfree)
    assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, Kt, lt)) < TOL
    assert rel(op.jacobian(op.allocate_jacobian()).data, O.jacobian_values(Kt, nzpos, rowval.size)) < TOL


def test_oseen_host_local_order(hip_ctx):
    """The host's local order (Gridap's tables, which the library cannot see): a 3 x 2 x 2 x 6 tet space whose cell
    dofs come in a shuffled P2 node order, a shuffled P1 order and a layout with the pressure dofs first and the
    velocity dofs node-major, handed over with its tables.  Pattern, nzpos (in the host's local order) and colours
    equal the oracle's on those same host-ordered ids bit for bit, and every nzpos entry addresses the CSC entry
    (row of test dof ii, column of trial dof jj); residual and CSC values equal the oracle's (1e-12), also with the
    step's point data made from FE functions and the outlet terms."""
    from oracle import oseen as O

    from hfk.tetmesh import outlet_faces, reorder_local

    sp = tet_p2p1_space((3, 2, 2), domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    rng = np.random.default_rng(23)
    nodeperm, pnodeperm = rng.permutation(10), rng.permutation(4)
    ldof = np.empty(34, dtype=np.int64)
    for n in range(10):
        for c in range(3):
            ldof[c * 10 + n] = 4 + 3 * n + c
    ldof[30:] = np.arange(4)
    hs = reorder_local(sp, nodeperm, pnodeperm, ldof)
    assert not np.array_equal(hs.cell_dofs, sp.cell_dofs)
    qp, qw = tet_quadrature_deg5()
    fields = rng.standard_normal((sp.ncells, len(qw), 9))
    dv = rng.standard_normal(sp.ndir)
    op = OseenOperator(hs, rho=1.0, mu=0.05, dt=0.01, bdf2=True, dirichlet_values=dv, ctx=hip_ctx)
    op.set_fields(fields[..., 0:3], fields[..., 3:6], fields[..., 6:9])
    colptr, rowval, nzpos = op.pattern()
    c2, r2, p2 = O.pattern(hs.cell_dofs, hs.nfree)
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    assert np.array_equal(op.colors(), O.greedy_colors(hs.cell_dofs, hs.nfree))
    nz = nzpos.reshape(sp.ncells, 34, 34, order="F")  # nz[cell, ii, jj] = nzpos[cell, ii + 34 jj]
    col_of = np.repeat(np.arange(colptr.size - 1), np.diff(colptr))
    for cell in range(0, sp.ncells, 7):
        ids = hs.cell_dofs[cell]
        for ii in range(34):
            for jj in range(34):
                p = nz[cell, ii, jj]
                if ids[ii] > 0 and ids[jj] > 0:
                    assert p > 0 and rowval[p - 1] == ids[ii] and col_of[p - 1] + 1 == ids[jj]
                else:
                    assert p == 0
    # values: the oracle in the library's own order (global quantities do not depend on the local order)
    _, _, p_std = O.pattern(sp.cell_dofs, sp.nfree)
    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0)
    A = op.jacobian(op.allocate_jacobian())
    assert rel(A.data, O.jacobian_values(K, p_std, r2.size)) < TOL
    x = rng.standard_normal(sp.nfree)
    assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, K, load)) < TOL
    # the step's point data from FE functions and the outlet terms, against the operator in the library's order
    ref = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.01, bdf2=True, dirichlet_values=dv, ctx=hip_ctx)
    of = outlet_faces(sp)
    un, unm1 = rng.standard_normal(sp.nfree), rng.standard_normal(sp.nfree)
    und, unm1d = rng.standard_normal(sp.ndir), rng.standard_normal(sp.ndir)
    for o in (op, ref):
        o.set_outlet(of, None, Pc=0.7, ZA=2.3)
        o.step_fields(un, und, unm1, unm1d)
    assert rel(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x)) < 1e-13
    assert rel(op.jacobian(op.allocate_jacobian()).data, ref.jacobian(ref.allocate_jacobian()).data) < 1e-13
    from copy import copy

    from hfk._lib import HFKError

    bad = copy(hs)
    bad.nodeperm = np.zeros(10, dtype=np.int32)  # not a permutation
    with pytest.raises(HFKError, match="ArgumentError"):
        OseenOperator(bad, ctx=hip_ctx)


def test_generated_point_form(hip_ctx):
    """The volume integrand compiled at run time from the captured form (hfk.formgen, hiprtc), in place of the
    built-in kernels: the reference's SUPG/PSPG/LSIC integrand gives the built-in operator's residual and CSC
    values and the oracle's (1e-12) with the Laplacian and the outlet terms, in the colour-ordered and the atomic
    launches, after update(rho, mu, dt) too; a stabilisation variant the library does not build in (the transient
    SUPG parameter) equals the oracle with that tau in numpy; removing the form restores the built-in kernels."""
    from oracle import oseen as O

    from hfk import formgen as fg
    from hfk.tetmesh import outlet_faces

    print("oseen generated point form", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, (3, 2, 2), True, seed=29)
    ref = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.01, bdf2=True, dirichlet_values=dv, ctx=hip_ctx)
    ref.set_fields(fields[..., 0:3], fields[..., 3:6], fields[..., 6:9])
    lap = rng.standard_normal((sp.ncells, len(qw), 3))
    of = outlet_faces(sp)
    wb = rng.standard_normal((len(of.cell), of.xb.shape[1], 3))
    for o in (op, ref):
        o.set_laplacian(lap)
        o.set_outlet(of, wb, Pc=0.7, ZA=2.3)
    op.set_form(fg.oseen_supg())
    colptr, rowval, nzpos = op.pattern()
    x = rng.standard_normal(sp.nfree)
    xb = np.concatenate([of.xb, wb], axis=2)
    Kf, Lf = O.outlet_terms(of.xo, xb, of.normal, 1.0, 0.7, 2.3)

    def oracle_check(rho, mu, c0, tau=None):
        K, load = O.element_system(sp.coords, qp, qw, fields, rho, mu, c0, lap=lap, tau=tau)
        Kf, Lf = O.outlet_terms(of.xo, xb, of.normal, rho, 0.7, 2.3)
        Kt, lt = O.add_faces(K, load, of.cell, Kf, Lf)
        A = op.jacobian(op.allocate_jacobian())
        assert rel(A.data, O.jacobian_values(Kt, nzpos, rowval.size)) < TOL
        assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, Kt, lt)) < TOL
        return A

    for det in (True, False):
        op.set_deterministic(det)
        ref.set_deterministic(det)
        A = oracle_check(1.0, 0.05, op.c0)
        assert rel(A.data, ref.jacobian(ref.allocate_jacobian()).data) < TOL
        assert rel(op.residual(np.zeros(sp.nfree), x), ref.residual(np.zeros(sp.nfree), x)) < TOL
    op.update(rho=1.3, mu=0.02, dt=0.02)
    oracle_check(1.3, 0.02, op.c0)
    # a stabilisation variant: the time scale under the root of tau_s
    ct = 0.5

    def tau_transient(D):
        unn = fg.norm(D.un)
        r0, r1, r2 = D.param("ct") * D.c0, 2.0 * unn / D.h, 3.0 * (4.0 * D.mu / (D.rho * D.h * D.h))
        ts = 1.0 / fg.sqrt(r0 * r0 + r1 * r1 + r2 * r2)
        return ts, ts / D.rho, D.rho * unn * D.h * 0.5

    def tau_np(unn, h, rho, mu, c0):
        ts = ((ct * c0) ** 2 + (2 * unn / h) ** 2 + 9 * (4 * mu / (rho * h ** 2)) ** 2) ** -0.5
        return ts, ts / rho, rho * unn * h * 0.5

    op.set_form(fg.oseen_supg(tau=tau_transient, params=("ct",), name="OseenTransientSupg"), {"ct": ct})
    oracle_check(1.3, 0.02, op.c0, tau=tau_np)
    op.set_form(None)
    oracle_check(1.3, 0.02, op.c0)


def test_field_block_sparsity(hip_ctx):
    """Field blocks (test/test_block_pattern.jl): the Oseen integrand without PSPG (generated form with tp = 0) has
    no pressure-pressure term, so its blocks leave that block out; the device pattern, nzpos and colours equal the
    oracle's block pattern bit for bit, the pressure-pressure part of the matrix is empty, and the values equal the
    oracle's (1e-12).  Then the block on half of the cells only (per-cell masks, as a coupling on part of the
    domain), and back to every block (the full pattern)."""
    from oracle import oseen as O

    from hfk import formgen as fg

    print("oseen field blocks", flush=True)
    sp, qp, qw, fields, dv, op, rng = setup(hip_ctx, (3, 2, 2), True, seed=31)

    def tau_no_pspg(D):
        ts, _, tl = fg.tau_reference(D)
        return ts, fg.const(0.0), tl

    form = fg.oseen_supg(tau=tau_no_pspg, name="OseenNoPspg")
    blocks = form.blocks()
    assert blocks.tolist() == [[True, True], [True, False]]
    op.set_form(form)
    op.set_blocks(blocks)
    fo = np.array([0] * 30 + [1] * 4)
    colptr, rowval, nzpos = op.pattern()
    cm = np.broadcast_to(blocks, (sp.ncells, 2, 2))
    c2, r2, p2 = O.block_pattern(sp.cell_dofs, sp.nfree, fo, cm)
    assert np.array_equal(colptr, c2) and np.array_equal(rowval, r2) and np.array_equal(nzpos, p2)
    assert rowval.size < O.pattern(sp.cell_dofs, sp.nfree)[1].size
    A = op.jacobian(op.allocate_jacobian())
    nu = sp.nfree_u
    assert A[nu:, nu:].nnz == 0

    def tau_np(unn, h, rho, mu, c0):
        ts = ((2 * unn / h) ** 2 + 9 * (4 * mu / (rho * h ** 2)) ** 2) ** -0.5
        return ts, 0.0 * ts, rho * unn * h * 0.5

    K, load = O.element_system(sp.coords, qp, qw, fields, 1.0, 0.05, op.c0, tau=tau_np)
    assert rel(A.data, O.jacobian_values(K, p2, r2.size)) < TOL
    x = rng.standard_normal(sp.nfree)
    assert rel(op.residual(np.zeros(sp.nfree), x), O.residual(sp, x, dv, K, load)) < TOL
    part = np.array(cm)
    part[rng.random(sp.ncells) < 0.5, 1, 1] = True
    op.set_blocks(cell_mask=part)
    colptr, rowval, nzpos = op.pattern()
    c3, r3, p3 = O.block_pattern(sp.cell_dofs, sp.nfree, fo, part)
    assert np.array_equal(colptr, c3) and np.array_equal(rowval, r3) and np.array_equal(nzpos, p3)
    A = op.jacobian(op.allocate_jacobian())
    assert rel(A.data, O.jacobian_values(K, p3, r3.size)) < TOL
    op.set_blocks(None)
    assert np.array_equal(op.pattern()[1], O.pattern(sp.cell_dofs, sp.nfree)[1])
