"""GPU parity of the assembled multi-field system (hfk_sys_*) on the Oseen P2-P1 Jacobian against the oracle
(oracle/blocks.py, pinned in tests/test_oracle_blocks.py) and scipy:

* y = A x through the CSR copy equals scipy's product (1e-13); the extracted blocks are the matrix entries bit for
  bit, their inverses the oracle's (B Binv = I, 1e-12); block application and damped sweeps equal the oracle's;
* GMRES and FGMRES preconditioned by the coupled velocity-pressure blocks (or sweeps of them) reach scipy's direct
  solution of the lifted step A x = -r(0) (1e-8);
* without PSPG (a compiled form with tau_p = 0) the pressure block is zero: the reference's per-field blocks leave
  every pressure block singular (counted; the strict mode raises), the coupled nodal blocks are all invertible;
  Vanka patches (a pressure with the velocity dofs of its vertex star) inverted on the device equal numpy's
  inverses of the extracted patch matrices, their restricted additive Schwarz equals the oracle's, and GMRES with
  it reaches the direct solution of the saddle point (1e-8), as it does with PSPG.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

from hfk._lib import HFKError
from hfk.assembled import AssembledSystem, taylor_hood_blocks, taylor_hood_vanka_patches
from hfk.oseen import OseenOperator
from hfk.tetmesh import tet_p2p1_space
from oracle import blocks as OB

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def oseen(hip_ctx, n=(4, 3, 3), form=None, seed=11):
    sp = tet_p2p1_space(n, domain=(0.0, 2.0, 0.0, 1.0, 0.0, 1.0))
    rng = np.random.default_rng(seed)
    from hfk.tetmesh import divfree_field

    op = OseenOperator(sp, rho=1.0, mu=0.05, dt=0.05, bdf2=True, dirichlet_values=rng.standard_normal(sp.ndir),
                       ctx=hip_ctx)
    w = divfree_field(3)
    op.set_steady_advection(w)
    if form is not None:
        op.set_form(form)
    A = op.jacobian(op.allocate_jacobian())
    b = -op.residual(np.zeros(sp.nfree), np.zeros(sp.nfree))
    return sp, op, A, b


def test_products_blocks_and_sweeps(hip_ctx):
    print("assembled system: products and blocks", flush=True)
    sp, op, A, b = oseen(hip_ctx)
    colptr, rowval, _ = op.pattern()
    S = AssembledSystem(colptr, rowval, ctx=hip_ctx)
    with pytest.raises(HFKError):  # used before its values are set
        S.matvec(np.ones(sp.nfree))
    blocks = taylor_hood_blocks(sp, "coupled")
    S.set_blocks(blocks)
    assert S.update(A.data) == 0
    rng = np.random.default_rng(1)
    x = rng.standard_normal(sp.nfree)
    assert rel(S.matvec(x), A @ x) < 1e-13
    B, Bi = S.blocks()
    Bo = OB.extract(A, blocks)
    assert np.array_equal(B, Bo)
    Bio, bad = OB.invert(Bo)
    assert bad == 0
    assert rel(Bi, Bio) < 1e-12
    assert max(np.abs(B[k] @ Bi[k] - np.eye(OB.BLOCK)).max() for k in range(len(B))) < 1e-10
    r = rng.standard_normal(sp.nfree)
    assert rel(S.apply_blocks(r), OB.apply(Bio, blocks, r)) < 1e-12
    assert rel(S.smooth(r, 3, 0.7), OB.sweeps(A, Bio, blocks, r, 3, 0.7)) < 1e-12
    with pytest.raises(HFKError):  # a dof in two blocks
        S.set_blocks(np.concatenate([blocks, blocks[:1]]))
    with pytest.raises(HFKError):  # a dof in no block
        S.set_blocks(blocks[1:])


@pytest.mark.parametrize("krylov,precond", [("gmres", "blocks"), ("fgmres", "blocks"), ("gmres", "sweeps")])
def test_solve_matches_direct(hip_ctx, krylov, precond):
    print(f"assembled system: {krylov} {precond}", flush=True)
    sp, op, A, b = oseen(hip_ctx)
    colptr, rowval, _ = op.pattern()
    S = AssembledSystem(colptr, rowval, ctx=hip_ctx).set_blocks(taylor_hood_blocks(sp, "coupled"))
    S.update(A.data)
    x, res = S.solve(b, krylov=krylov, preconditioner=precond, sweeps=2, omega=0.8, rtol=1e-10, maxiter=3000,
                     restart=60)
    xd = sla.spsolve(A.tocsc(), b)
    assert res.converged and res.residual <= 1e-10 * res.rhs_norm * 1.0001
    assert rel(x, xd) < 1e-8
    assert np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
    print(f"{krylov} {precond}: {res.iterations} iterations", flush=True)


def test_saddle_point_needs_coupled_blocks(hip_ctx):
    from hfk import formgen as fg

    print("assembled system: saddle point without PSPG", flush=True)

    def tau_no_pspg(D):
        ts, _, tl = fg.tau_reference(D)
        return ts, fg.const(0.0), tl

    sp, op, A, b = oseen(hip_ctx, form=fg.oseen_supg(tau=tau_no_pspg, name="OseenNoPspg"))
    nu = sp.nfree_u
    assert np.abs(A[nu:, nu:]).max() == 0.0  # no pressure diagonal
    colptr, rowval, _ = op.pattern()
    S = AssembledSystem(colptr, rowval, ctx=hip_ctx)
    fields = taylor_hood_blocks(sp, "fields")
    S.set_blocks(fields)
    npres = sp.vert_xyz.shape[0]
    assert S.update(A.data) == npres  # every pressure block is singular
    assert OB.invert(OB.extract(A, fields))[1] == npres
    with pytest.raises(np.linalg.LinAlgError):
        S.update(A.data, strict=True)
    S.set_blocks(taylor_hood_blocks(sp, "coupled"))
    assert S.update(A.data) == 0  # every coupled block is invertible


def no_pspg_form():
    from hfk import formgen as fg

    def tau_no_pspg(D):
        ts, _, tl = fg.tau_reference(D)
        return ts, fg.const(0.0), tl

    return fg.oseen_supg(tau=tau_no_pspg, name="OseenNoPspg")


@pytest.mark.parametrize("pspg", [False, True])
def test_vanka_patches(hip_ctx, pspg):
    print(f"assembled system: Vanka patches, pspg={pspg}", flush=True)
    sp, op, A, b = oseen(hip_ctx, form=None if pspg else no_pspg_form())
    colptr, rowval, _ = op.pattern()
    S = AssembledSystem(colptr, rowval, ctx=hip_ctx)
    ptr, dofs, owned = taylor_hood_vanka_patches(sp)
    S.set_patches(ptr, dofs, owned)
    assert S.update(A.data) == 0
    inv = S.patch_inverses()
    Ad = A.tocsc()
    worst = 0.0
    for k in range(len(ptr) - 1):
        ix = dofs[ptr[k]:ptr[k + 1]] - 1
        Pk = Ad[ix][:, ix].toarray()
        worst = max(worst, rel(inv[k], np.linalg.inv(Pk)))
    assert worst < 1e-10
    rng = np.random.default_rng(4)
    r = rng.standard_normal(sp.nfree)
    pats = [(list(dofs[ptr[k]:ptr[k + 1]]), list(owned[ptr[k]:ptr[k + 1]])) for k in range(len(ptr) - 1)]
    assert rel(S.apply_patches(r), OB.ras_apply(A, pats, r)) < 1e-10
    x, res = S.solve(b, krylov="gmres", preconditioner="patches", rtol=1e-10, maxiter=2000, restart=60)
    assert res.converged
    assert rel(x, sla.spsolve(Ad, b)) < 1e-8
    print(f"Vanka pspg={pspg}: {res.iterations} GMRES iterations", flush=True)
    own2 = owned.copy()
    own2[np.nonzero(owned == 0)[0][0]] = 1
    with pytest.raises(HFKError):  # a dof owned by two patches
        S.set_patches(ptr, dofs, own2)
