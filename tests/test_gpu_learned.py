"""The learned coarse correction on the GPU.

* forward output, loss and weight gradients of the device U-Net (BF16 MFMA kernels) against a torch FP32 CPU
  restatement with the same BF16 rounding points (tests/learned_ref.py);
* training with the device kernels reduces the loss on pairs produced by the plain p-multigrid;
* inside the p-multigrid the network (as the coarse PCG's initial guess, or as the whole coarse correction) leaves
  the answer to the Krylov tolerance: the same solution as the plain cycle.
"""
import numpy as np
import pytest

from hfk import HyperelasticForm, KernelOperator, MatrixFreeJacobian
from hfk.learned import CoarseNetwork, coarse_pairs_dev, set_coarse_network
from hfk.mesh import dirichlet_closed_x0, hex_space
from hfk.solvers import Chebyshev, MatrixFreeSolver, PMultigrid
from hfk.synthetic import smooth_field

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300)


def test_network_gradient_matches_torch(hip_ctx):
    import learned_ref

    cs = hex_space((16, 16, 16), 1, dirichlet=dirichlet_closed_x0)
    net = CoarseNetwork(hip_ctx, cs, seed=3)
    w = net.weights()
    rng = np.random.default_rng(0)
    r = rng.standard_normal(cs.nfree)
    e = rng.standard_normal(cs.nfree)
    loss, g = net.gradient(r, e)
    Lr, gr, out = learned_ref.loss_and_grad(w, cs.lattice_dofs(), 17, 1.0 / 16, r, e)
    print(f"loss gpu {loss:.6e} torch {Lr:.6e}; grad rel {rel(g, gr):.3e}", flush=True)
    assert loss == pytest.approx(Lr, rel=2e-2)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.995 and rel(g, gr) < 0.08
    # the forward output through apply: z = (s/h) out on the free dofs
    z = net.apply(r)
    s = np.sqrt(r @ r / r.size)
    ld = cs.lattice_dofs()
    zr = np.zeros(cs.nfree)
    for c in range(3):
        f = ld[:, c] > 0
        zr[ld[f, c] - 1] = out[c, f] * s * 16
    assert rel(z, zr) < 2e-2


def make_problem(hip_ctx, n, seed=0, mu_sigma=0.5):
    from hfk.synthetic import lognormal_field

    fs = hex_space((n, n, n), 2, dirichlet=dirichlet_closed_x0)
    cs = hex_space((n, n, n), 1, dirichlet=dirichlet_closed_x0)
    u = smooth_field(seed + 2)
    form = HyperelasticForm("neohookean", degree=4)
    fop = KernelOperator(form, fs, jacobian="matrixfree", dirichlet_values=fs.dirichlet_values(u), ctx=hip_ctx)
    cop = KernelOperator(form, cs, jacobian="assembled", ctx=hip_ctx)
    # lognormal shear modulus at the quadrature points (both levels use the 27-point rule)
    from oracle.oracle import gauss01

    g, _ = gauss01(3)
    pts = np.array([[g[a], g[b], g[c]] for c in range(3) for b in range(3) for a in range(3)])
    mu = lognormal_field(seed + 40, sigma=mu_sigma, ell=0.2)
    lo = fs.coords[:, 0, :]
    hi = fs.coords[:, 7, :]
    xq = lo[:, None, :] + pts[None, :, :] * (hi - lo)[:, None, :]
    mup = mu(xq.reshape(-1, 3))
    fop.set_mu_field(mup)
    cop.set_mu_field(mup)
    mf = MatrixFreeJacobian(fop).prepare(fs.free_values(u))
    return fs, cs, fop, cop, mf


def test_training_and_learned_cycle(hip_ctx):
    import torch

    n = 16
    fs, cs, fop, cop, mf = make_problem(hip_ctx, n)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-2).update()
    assert pmg.ok
    dev = torch.device("cuda:0")
    ns = 48
    rng = np.random.default_rng(1)
    B = np.stack([smooth_field(100 + k, amp=1.0)(fs.node_xyz[fs.free_node_of]).reshape(-1)
                  + 0.1 * rng.standard_normal(fs.nfree) for k in range(ns)])
    bd = torch.from_numpy(B).to(dev)
    bc = torch.zeros((ns, cs.nfree), dtype=torch.float64, device=dev)
    ec = torch.zeros_like(bc)
    torch.cuda.synchronize()
    coarse_pairs_dev(pmg, bd.data_ptr(), ns, bc.data_ptr(), ec.data_ptr(), rtol=1e-10)
    net = CoarseNetwork(hip_ctx, cs, seed=7)
    hist = net.train_dev(bc.data_ptr(), ec.data_ptr(), ns - 8, steps=300, batch=4, lr=1e-3)
    print("loss first/last:", hist[:5].mean(), hist[-20:].mean(), flush=True)
    assert hist[-20:].mean() < 0.5 * hist[:5].mean()
    # held-out pairs: the network's relative error
    bch = bc[ns - 8:].cpu().numpy()
    ech = ec[ns - 8:].cpu().numpy()
    errs = [rel(net.apply(bch[k]), ech[k]) for k in range(8)]
    print("held-out relative error of the coarse correction:", np.mean(errs), flush=True)
    # the cycle with the network: same answer (the Krylov tolerance decides)
    b = rng.standard_normal(fs.nfree)
    sols = {}
    for mode in ("off", "warmstart", "replace"):
        set_coarse_network(pmg, None if mode == "off" else net, mode)
        x = np.zeros(fs.nfree)
        info0 = pmg.info()["coarse_iters"]
        res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=200, preconditioner=pmg).solve(x, mf, b)
        inner = pmg.info()["coarse_iters"] - info0
        print(f"mode {mode}: {res}, coarse iterations {inner}", flush=True)
        assert res.converged
        sols[mode] = (x, res.iterations, inner)
    assert rel(sols["warmstart"][0], sols["off"][0]) < 1e-8
    assert rel(sols["replace"][0], sols["off"][0]) < 1e-8
    assert sols["warmstart"][2] < sols["off"][2]
