"""The learned coarse correction on the GPU.

* forward output, loss and weight gradients of the device U-Net (BF16 MFMA kernels) against a torch FP32 CPU
  restatement with the same BF16 rounding points (tests/learned_ref.py);
* energy training with the device kernels on residuals produced by the plain p-multigrid lowers the held-out
  energy error of the coarse correction;
* inside the p-multigrid, every way of using the network leaves the answer to the Krylov tolerance (the plain
  cycle's solution), and as the preconditioner of the coarse flexible CG it cuts the coarse iterations without
  adding outer ones.
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


@pytest.mark.parametrize("coarsen", [False, True])
def test_network_gradient_matches_torch(hip_ctx, coarsen):
    import learned_ref

    cs = hex_space((16, 16, 16), 1, dirichlet=dirichlet_closed_x0)
    net = CoarseNetwork(hip_ctx, cs, seed=3, coarsen=coarsen)
    w = net.weights()
    rng = np.random.default_rng(0)
    r = rng.standard_normal(cs.nfree)
    e = rng.standard_normal(cs.nfree)
    loss, g = net.gradient(r, e)
    # the weight gradients are summed in a fixed order: the same gradient bit for bit on a second evaluation
    loss2, g2 = net.gradient(r, e)
    assert loss2 == loss and np.array_equal(g2, g)
    Lr, gr, out = learned_ref.loss_and_grad(w, cs.lattice_dofs(), 17, 1.0 / 16, r, e, coarsen=coarsen)
    print(f"loss gpu {loss:.6e} torch {Lr:.6e}; grad rel {rel(g, gr):.3e}", flush=True)
    # BF16 activations and weights with FP32 accumulation on both sides; the two differ only by summation order
    # (measured: loss 2e-5, weight gradient 1.7e-3 to 1.8e-3 relative)
    assert loss == pytest.approx(Lr, rel=1e-3)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.99998 and rel(g, gr) < 5e-3
    # the forward output through apply: z = (s/h) out on the free dofs
    z = net.apply(r)
    s = np.sqrt(r @ r / r.size)
    ld = cs.lattice_dofs()
    zr = np.zeros(cs.nfree)
    for c in range(3):
        f = ld[:, c] > 0
        zr[ld[f, c] - 1] = out[c, f] * s * 16
    assert rel(z, zr) < 2e-2


@pytest.mark.parametrize("coarsen", [False, True])
def test_energy_loss_gradient_matches_torch(hip_ctx, oracle, coarsen):
    """The training objective itself: the energy loss L = (w'(A/h)w - 2 w'x)/n (x = r/s) of the device training,
    summed over three coarse samples, and its weight gradient, against the torch restatement with the coarse matrix
    of the oracle (the hierarchy at the zero state: the coarse matrix is the oracle's Jacobian at zero with the
    fine rule)."""
    import torch

    import learned_ref
    from hfk.learned import energy_gradient_dev

    fs = hex_space((16, 16, 16), 2, dirichlet=dirichlet_closed_x0)
    cs = hex_space((16, 16, 16), 1, dirichlet=dirichlet_closed_x0)
    form = HyperelasticForm("neohookean", degree=4)
    fop = KernelOperator(form, fs, jacobian="matrixfree", ctx=hip_ctx)
    cop = KernelOperator(form, cs, jacobian="assembled", ctx=hip_ctx)
    MatrixFreeJacobian(fop).prepare(np.zeros(fs.nfree))
    pmg = PMultigrid([fop, cop], Chebyshev(degree=3)).update()
    net = CoarseNetwork(hip_ctx, cs, seed=5, coarsen=coarsen)
    set_coarse_network(pmg, net, "cycle", scaled=False)
    rng = np.random.default_rng(2)
    bc = rng.standard_normal((3, cs.nfree))
    bcd = torch.tensor(bc, device="cuda")
    torch.cuda.synchronize()
    loss, g = energy_gradient_dev(pmg, net, bcd.data_ptr(), 3)
    loss2, g2 = energy_gradient_dev(pmg, net, bcd.data_ptr(), 3)
    assert loss2 == loss and np.array_equal(g2, g)  # fixed-order reductions
    A = oracle.OracleOp(cs, "neohookean", 2.0, 1.0, points1d=3).csc(np.zeros(cs.nfree))
    w = net.weights()
    Lr, gr = 0.0, np.zeros_like(g, dtype=np.float64)
    for k in range(3):
        lk, gk = learned_ref.energy_loss_and_grad(w, cs.lattice_dofs(), 17, 1.0 / 16, A, bc[k], coarsen=coarsen)
        Lr += lk
        gr += gk
    print(f"energy loss gpu {loss:.8e} torch {Lr:.8e}; grad rel {rel(g, gr):.3e}", flush=True)
    assert loss == pytest.approx(Lr, rel=1e-3)
    cos = float(np.dot(g, gr) / (np.linalg.norm(g) * np.linalg.norm(gr)))
    assert cos > 0.99998 and rel(g, gr) < 5e-3


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


@pytest.mark.parametrize("scaled", [False, True])
def test_training_and_learned_cycle(hip_ctx, scaled):
    """Energy training on restricted, coarse-smoothed p-MG residuals; then the network inside the cycle: every mode
    reaches the plain cycle's answer (the outer tolerance decides), and as the preconditioner of the coarse flexible
    CG it cuts the coarse iterations without adding outer ones -- the gain of the trained network."""
    import torch

    from hfk.learned import coarse_error_dev, train_energy_dev

    n = 32
    fs, cs, fop, cop, mf = make_problem(hip_ctx, n)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-2, coarse_maxiter=400).update()
    assert pmg.ok
    dev = torch.device("cuda:0")
    ns, nh = 160, 16
    gen = torch.Generator(device=dev).manual_seed(1)
    bd = torch.randn((ns, fs.nfree), dtype=torch.float64, device=dev, generator=gen)
    bc = torch.zeros((ns, cs.nfree), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    coarse_pairs_dev(pmg, bd.data_ptr(), ns, bc.data_ptr(), None, coarse_smoothed=True)
    del bd
    set_coarse_network(pmg, None, "off", 0, scaled)
    net = CoarseNetwork(hip_ctx, cs, seed=7, coarsen=True)
    err0 = coarse_error_dev(pmg, net, bc[ns - nh:].data_ptr(), nh)
    hist = train_energy_dev(pmg, net, bc.data_ptr(), ns - nh, steps=500, batch=4, lr=1e-3)
    err1 = coarse_error_dev(pmg, net, bc[ns - nh:].data_ptr(), nh)
    print(f"energy loss first/last: {hist[:5].mean():.4f} {hist[-20:].mean():.4f}; held-out energy error "
          f"{err0:.3f} -> {err1:.3f}", flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-20:].mean() < hist[:5].mean() - 1.0
    assert err1 < 0.85 * err0
    rng = np.random.default_rng(1)
    b = rng.standard_normal(fs.nfree)
    sols = {}
    for mode in ("off", "warmstart", "cycle", "precondition"):
        set_coarse_network(pmg, None if mode == "off" else net, mode, 0, scaled)
        x = np.zeros(fs.nfree)
        info0 = pmg.info()["coarse_iters"]
        res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=400, preconditioner=pmg).solve(x, mf, b)
        inner = pmg.info()["coarse_iters"] - info0
        print(f"mode {mode}: {res}, coarse iterations {inner}", flush=True)
        assert res.converged
        sols[mode] = (x, res.iterations, inner)
    for mode in ("warmstart", "cycle", "precondition"):
        assert rel(sols[mode][0], sols["off"][0]) < 1e-7
    # the trained network as the coarse preconditioner: the outer iterations stay, the coarse work collapses
    assert sols["precondition"][1] <= sols["off"][1] + 1
    assert sols["precondition"][2] < 0.25 * sols["off"][2]
    # the network cycle inside restarted flexible GMRES (right preconditioned, the preconditioner may vary)
    set_coarse_network(pmg, net, "cycle", 0, scaled)
    x = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fgmres", rtol=1e-10, maxiter=400, preconditioner=pmg, restart=30).solve(x, mf, b)
    print(f"fgmres + network cycle: {res}", flush=True)
    assert res.converged and rel(x, sols["off"][0]) < 1e-7
    set_coarse_network(pmg, None, "off")


def test_fallback_to_plain_cycle(hip_ctx):
    """Plain p-MG is the fallback of the learned solve: a network whose output is not finite sets itself aside at
    its first direction (FCG and FGMRES), and a stagnation watch that fires at once (forced from the host) hands the
    solve to the plain cycle; both still reach the plain cycle's answer.  With the watch off the same network fails
    the solve."""
    fs, cs, fop, cop, mf = make_problem(hip_ctx, 16)
    pmg = PMultigrid([fop, cop], Chebyshev(degree=4), coarse_rtol=1e-2, coarse_maxiter=400).update()
    assert pmg.ok
    b = np.random.default_rng(3).standard_normal(fs.nfree)
    x_ref = np.zeros(fs.nfree)
    ref = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=400, preconditioner=pmg).solve(x_ref, mf, b)
    assert ref.converged and not ref.fell_back
    net = CoarseNetwork(hip_ctx, cs, seed=3, coarsen=True)
    good = net.weights()
    net.set_weights(np.full_like(good, np.nan))
    set_coarse_network(pmg, net, "cycle", 0, True)
    count = pmg.fallbacks()
    for krylov in ("fcg", "fgmres"):
        x = np.zeros(fs.nfree)
        res = MatrixFreeSolver(krylov, rtol=1e-10, maxiter=400, preconditioner=pmg, restart=30).solve(x, mf, b)
        print(f"non-finite network, {krylov}: {res}", flush=True)
        assert res.converged and res.fell_back
        assert rel(x, x_ref) < 1e-7
        count += 1
        assert pmg.fallbacks() == count
    pmg.set_fallback(window=0)
    x = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=400, preconditioner=pmg).solve(x, mf, b)
    assert not res.converged and not res.fell_back
    # an untrained network with a watch that fires at the second iteration
    net.set_weights(good)
    pmg.set_fallback(window=1, factor=1e-300)
    x = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=400, preconditioner=pmg).solve(x, mf, b)
    print(f"forced stagnation: {res}", flush=True)
    assert res.converged and res.fell_back and rel(x, x_ref) < 1e-7
    # the default watch: an untrained network that still converges keeps it or falls back, the answer is the same
    pmg.set_fallback()
    x = np.zeros(fs.nfree)
    res = MatrixFreeSolver("fcg", rtol=1e-10, maxiter=400, preconditioner=pmg).solve(x, mf, b)
    assert res.converged and rel(x, x_ref) < 1e-7
    with pytest.raises(RuntimeError):
        pmg.set_fallback(window=4, factor=1.5)
    set_coarse_network(pmg, None, "off")
