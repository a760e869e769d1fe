"""Pins of the oracle's solver-layer restatements (CPU, small problems).

* nodal blocks from the stored tangents = the 3x3 diagonal blocks of the assembled matrix
  (test/solvers/test_chebyshev.jl:40-62);
* Chebyshev: the damping factor's known answer 1/cosh(3 acosh(21/19)) (test_chebyshev.jl:123), and every
  eigencomponent of M^-1 A inside [a, b] is damped by at most 1/T_k(theta/delta) (test_chebyshev.jl:64-98);
* Arnoldi bounds: Ritz values inside the spectrum, the largest within 10 % of lambda_max after 10 steps
  (test_matrixfree_solver.jl);
* prolongation: exact on linear fields, rows of E sum to 1 (partition of unity).
"""
import numpy as np
import pytest
import scipy.linalg as sla

from hfk.mesh import dirichlet_closed_x0, hex_space
from hfk.solvers import Chebyshev


def small_problem(oracle, n=(3, 2, 2), order=2, law="neohookean", seed=0):
    sp = hex_space(n, order, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    x = 0.01 / order * rng.standard_normal(sp.nfree)
    dv = 0.01 / order * rng.standard_normal(sp.ndir)
    op = oracle.OracleOp(sp, law, 2.0, 1.0, dvals=dv)
    return sp, op, x


def test_blocks_are_assembled_diagonal(oracle):
    sp, op, x = small_problem(oracle)
    dofs, Binv = op.blocks_from_tangents(op.prepare(x))
    K = op.csc(x).toarray()
    assert dofs.shape[0] == sp.nfree // 3  # every node has 3 free components (closed face: all fixed)
    for k in range(dofs.shape[0]):
        d = dofs[k] - 1
        B = K[np.ix_(d, d)]
        assert np.allclose(Binv[k].reshape(3, 3, order="F"), np.linalg.inv(B), rtol=1e-10, atol=1e-12)


def test_chebyshev_damping_known_answer():
    assert Chebyshev(degree=3, range=20.0).damping() == pytest.approx(1 / np.cosh(3 * np.arccosh(21 / 19)), rel=1e-14)


@pytest.mark.parametrize("degree", [2, 3, 4])
def test_chebyshev_damps_every_component(oracle, degree):
    sp, op, x = small_problem(oracle, n=(2, 2, 1))
    A = op.csc(x).toarray()
    dofs, Binv = op.blocks_from_tangents(op.prepare(x))
    Minv = oracle.block_inverse_matrix(dofs, Binv, sp.nfree).toarray()
    lam, W = sla.eig(Minv @ A)
    lam, W = lam.real, W.real
    b = 1.0 * lam.max()
    a = b / 20.0
    cfg = Chebyshev(degree=degree, range=20.0)
    bound = cfg.damping()
    for i in np.argsort(lam)[::-1][:40]:
        if lam[i] < a:
            continue
        e0 = W[:, i]
        # error propagation of the smoother on A u = 0 from u = e0: u_k = p_k(M^-1 A) e0
        uk = oracle.chebyshev_smooth(A, Minv, e0, np.zeros(sp.nfree), a, b, degree)
        assert np.linalg.norm(uk) <= (bound + 1e-8) * np.linalg.norm(e0)


def test_arnoldi_bounds_inside_spectrum(oracle):
    sp, op, x = small_problem(oracle, n=(2, 2, 2))
    A = op.csc(x).toarray()
    dofs, Binv = op.blocks_from_tangents(op.prepare(x))
    Minv = oracle.block_inverse_matrix(dofs, Binv, sp.nfree).toarray()
    ev = np.linalg.eigvals(A @ Minv).real
    lo, hi = oracle.arnoldi_bounds(lambda v: A @ (Minv @ v), oracle.chebyshev_start(sp.nfree), 10)
    assert ev.min() - 1e-10 <= lo <= hi <= ev.max() + 1e-10
    assert hi >= 0.9 * ev.max()


def test_chebyshev_start_formula(oracle):
    v = oracle.chebyshev_start(5)
    i = np.arange(1, 6)
    w = np.sin(1 + 0.7 * i) + 0.3 * np.cos(2.3 * i)
    assert np.allclose(v, w / np.linalg.norm(w), atol=1e-16)


@pytest.mark.parametrize("qc,qf", [(1, 2), (2, 3), (1, 3)])
def test_prolongation_exact_on_linear_fields(oracle, qc, qf):
    E = oracle.lagrange_embedding(qc + 1, qf + 1)
    assert np.allclose(E.sum(axis=1), 1.0, atol=1e-13)
    cs = hex_space((2, 2, 1), qc, dirichlet=dirichlet_closed_x0, perturb=0.0)
    fs = hex_space((2, 2, 1), qf, dirichlet=dirichlet_closed_x0, perturb=0.0)
    P = oracle.prolongation_matrix(cs, fs)
    A = np.array([[0.3, -0.1, 0.2], [0.05, 0.4, 0.0], [0.1, 0.2, -0.3]])
    f = lambda xyz: xyz[:, :1] * (A[:, 0][None, :] + 0.0 * xyz)  # linear, zero on the clamped face x = 0
    assert np.allclose(P @ cs.free_values(f), fs.free_values(f), atol=1e-13)


def test_gmres_restatement_solves_and_restarts():
    """The oracle's restarted right-preconditioned GMRES: on a nonsymmetric system it reaches the tolerance on the
    true residual, with and without restarts, and with a Jacobi preconditioner; unrestarted GMRES terminates in at
    most n steps on an n x n system (the Krylov space is exhausted)."""
    from oracle.oracle import gmres_right

    rng = np.random.default_rng(4)
    n = 40
    A = np.eye(n) * 4 + rng.standard_normal((n, n)) * 0.3
    b = rng.standard_normal(n)
    d = np.diag(A).copy()
    for restart in (5, 50):
        for Minv in (lambda v: v, lambda v: v / d):
            x, it, ok = gmres_right(A, b, Minv, 1e-10, restart, 500)
            assert ok and np.linalg.norm(b - A @ x) <= 1.01e-10 * np.linalg.norm(b)
    x, it, ok = gmres_right(A, b, lambda v: v, 1e-14, n, 500)
    assert ok and it <= n
