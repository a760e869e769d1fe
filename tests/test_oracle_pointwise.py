"""Pins of the oracle's pointwise pieces against closed forms, textbook routines and finite differences.

* Gauss--Legendre on [0, 1]: the 3-point closed form (1/2 ± sqrt(3/5)/2, weights 5/18, 4/9, 5/18) and
  exactness up to degree 2P - 1.
* 1D Lagrange (src/tensorproduct.jl:24-47): the reference's known answers lagrange1d(3, 0.5) = [0, 1, 0],
  partition of unity and derivative sum 0 (test/test_matrixfree.jl:9-17), derivatives vs finite differences.
* Physical points of the tensor Gauss points: affine images on an affine cell (x fastest), the vertex mean for
  the one-point rule.
* Laws: P = dW/dF and T = dP/dF by central differences of the energy, the SVK closed forms of
  test/common.jl:2-8 written with numpy matrices, and C_svk at F = I (test/common.jl:11-20).
"""
import numpy as np
import pytest


def test_gauss_closed_form(oracle):
    x, w = oracle.gauss01(3)
    r = np.sqrt(3 / 5) / 2
    assert np.allclose(x, [0.5 - r, 0.5, 0.5 + r], atol=1e-15, rtol=0)
    assert np.allclose(w, [5 / 18, 4 / 9, 5 / 18], atol=1e-15, rtol=0)
    x, w = oracle.gauss01(1)
    assert x[0] == pytest.approx(0.5, abs=1e-15) and w[0] == pytest.approx(1.0, abs=1e-15)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
def test_gauss_exactness(oracle, P):
    x, w = oracle.gauss01(P)
    assert np.all(np.diff(x) > 0)
    for deg in range(2 * P):
        assert np.dot(w, x ** deg) == pytest.approx(1.0 / (deg + 1), abs=1e-14)
    # and not exact one degree higher (a wrong rule that happens to be exact further would be suspicious)
    assert abs(np.dot(w, x ** (2 * P)) - 1.0 / (2 * P + 1)) > 1e-8


def test_physical_points_closed_forms(oracle):
    """An affine (sheared, scaled) cell maps the Gauss points affinely, x fastest; the one-point rule of any
    trilinear cell is the mean of its vertices."""
    A = np.array([[0.5, 0.1, 0.0], [0.2, 0.3, -0.1], [0.0, 0.05, 0.7]])
    b = np.array([1.0, -2.0, 0.5])
    V = np.array([[v & 1, (v >> 1) & 1, (v >> 2) & 1] for v in range(8)], dtype=float)
    coords = (V @ A.T + b)[None]
    g, _ = oracle.gauss01(3)
    xi = np.array([[g[a], g[bb], g[c]] for c in range(3) for bb in range(3) for a in range(3)])
    assert np.abs(oracle.physical_points(coords, 3)[0] - (xi @ A.T + b)).max() < 1e-15
    X = np.random.default_rng(2).standard_normal((4, 8, 3))
    assert np.abs(oracle.physical_points(X, 1)[:, 0] - X.mean(axis=1)).max() < 1e-15


def test_lagrange_known_answers(oracle):
    N, _ = oracle.lagrange1d(3, 0.5)
    assert np.allclose(N, [0, 1, 0], atol=1e-15)
    for Q in (1, 2, 3, 4):
        for xi in (0.0, 0.13, 0.5, 0.77, 1.0):
            N, dN = oracle.lagrange1d(Q, xi)
            assert N.sum() == pytest.approx(1.0, abs=1e-14)
            assert dN.sum() == pytest.approx(0.0, abs=1e-12)
            h = 1e-6
            Np, _ = oracle.lagrange1d(Q, xi + h)
            Nm, _ = oracle.lagrange1d(Q, xi - h)
            assert np.allclose(dN, (Np - Nm) / (2 * h), atol=1e-7)
        if Q > 1:  # interpolation property at the nodes
            for i in range(Q):
                N, _ = oracle.lagrange1d(Q, i / (Q - 1))
                assert np.allclose(N, np.eye(Q)[i], atol=1e-14)


def _rand_g(rng, scale=0.1):
    return scale * rng.standard_normal(9)


def _F(g):
    # F[c, J] = delta + g[J + 3c]
    return np.eye(3) + g.reshape(3, 3)


@pytest.mark.parametrize("law", ["svk", "neohookean", "linear"])
def test_stress_is_energy_gradient(oracle, law):
    rng = np.random.default_rng(1)
    for _ in range(5):
        g = _rand_g(rng)
        R = oracle.law_P(law, 2.0, 1.0, g)
        h = 1e-6
        fd = np.zeros(9)
        for s in range(9):
            e = np.zeros(9)
            e[s] = h
            fd[s] = (oracle.law_W(law, 2.0, 1.0, g + e) - oracle.law_W(law, 2.0, 1.0, g - e)) / (2 * h)
        assert np.allclose(R, fd, atol=1e-8, rtol=1e-7)


@pytest.mark.parametrize("law", ["svk", "neohookean", "linear"])
def test_tangent_is_stress_derivative(oracle, law):
    rng = np.random.default_rng(2)
    for _ in range(5):
        g = _rand_g(rng)
        T = oracle.law_T(law, 2.0, 1.3, g).reshape(9, 9, order="F")  # T[sv, su]
        h = 1e-6
        for su in range(9):
            e = np.zeros(9)
            e[su] = h
            fd = (oracle.law_P(law, 2.0, 1.3, g + e) - oracle.law_P(law, 2.0, 1.3, g - e)) / (2 * h)
            assert np.allclose(T[:, su], fd, atol=1e-8, rtol=1e-7)
        assert np.allclose(T, T.T, atol=1e-13)  # hyperelastic: symmetric


def test_svk_closed_form(oracle):
    """test/common.jl:2-8 written with numpy matrices: P = F S, S = lam tr(E) I + 2 mu E, E = (F'F - I)/2."""
    rng = np.random.default_rng(3)
    lam, mu = 2.0, 1.0
    for _ in range(5):
        g = _rand_g(rng, 0.2)
        F = _F(g)
        E = 0.5 * (F.T @ F - np.eye(3))
        S = lam * np.trace(E) * np.eye(3) + 2 * mu * E
        Pm = F @ S
        R = oracle.law_P("svk", lam, mu, g)
        assert np.allclose(R.reshape(3, 3), Pm, atol=1e-14)  # R[J + 3c] = P[c, J]
        # dP(F, dF) = dF S + F dS(F, dF) against T applied to a direction
        dG = rng.standard_normal(9)
        dF = dG.reshape(3, 3)
        dE = 0.5 * (dF.T @ F + F.T @ dF)
        dS = lam * np.trace(dE) * np.eye(3) + 2 * mu * dE
        dP = dF @ S + F @ dS
        T = oracle.law_T("svk", lam, mu, g).reshape(9, 9, order="F")
        assert np.allclose((T @ dG).reshape(3, 3), dP, atol=1e-13)


@pytest.mark.parametrize("law", ["svk", "neohookean", "linear"])
def test_tangent_at_identity_is_C_svk(oracle, law):
    """At F = I every law here linearises to C[(iJ),(kL)] = lam d_iJ d_kL + mu (d_ik d_JL + d_iL d_Jk)
    (test/common.jl:11-20 with lam = 2, mu = 1)."""
    lam, mu = 2.0, 1.0
    T = oracle.law_T(law, lam, mu, np.zeros(9))
    d = np.eye(3)
    for c in range(3):
        for J in range(3):
            for k in range(3):
                for L in range(3):
                    ref = lam * d[c, J] * d[k, L] + mu * (d[c, k] * d[J, L] + d[c, L] * d[J, k])
                    assert T[(J + 3 * c) + 9 * (L + 3 * k)] == pytest.approx(ref, abs=1e-14)


def test_neohookean_closed_form(oracle):
    """P = mu (F - F^-T) + lam ln(J) F^-T for W = mu/2 (I1-3) - mu lnJ + lam/2 lnJ^2."""
    rng = np.random.default_rng(4)
    for _ in range(5):
        g = _rand_g(rng, 0.2)
        F = _F(g)
        Fit = np.linalg.inv(F).T
        Pm = 1.0 * (F - Fit) + 2.0 * np.log(np.linalg.det(F)) * Fit
        assert np.allclose(oracle.law_P("neohookean", 2.0, 1.0, g).reshape(3, 3), Pm, atol=1e-13)


@pytest.mark.parametrize("law", ["svk", "neohookean", "linear"])
def test_energy_vanishes_at_identity(oracle, law):
    """The absolute energy constant: W(F = I) = 0 for every law (the stress-free reference state), and the
    hyperelastic laws are objective: W(R) = 0 for a rigid rotation R."""
    assert oracle.law_W(law, 2.0, 1.0, np.zeros(9)) == 0.0
    if law != "linear":
        th = 0.7
        R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
        g = (R - np.eye(3)).reshape(-1)  # g[J + 3c] = F[c, J] - delta
        assert abs(oracle.law_W(law, 2.0, 1.0, g)) < 1e-14


def test_energy_closed_forms(oracle):
    """W of St Venant-Kirchhoff, lam/2 tr(E)^2 + mu E:E (test/common.jl:2-8 integrated), of the neo-Hookean law
    mu/2 (I1 - 3) - mu ln J + lam/2 (ln J)^2 and of the linear law lam/2 tr(eps)^2 + mu eps:eps, in numpy."""
    rng = np.random.default_rng(5)
    lam, mu = 2.0, 1.3
    for _ in range(5):
        g = _rand_g(rng, 0.2)
        F = _F(g)
        E = 0.5 * (F.T @ F - np.eye(3))
        assert oracle.law_W("svk", lam, mu, g) == pytest.approx(0.5 * lam * np.trace(E) ** 2 + mu * np.sum(E * E),
                                                                rel=1e-13)
        J = np.linalg.det(F)
        wn = 0.5 * mu * (np.sum(F * F) - 3) - mu * np.log(J) + 0.5 * lam * np.log(J) ** 2
        assert oracle.law_W("neohookean", lam, mu, g) == pytest.approx(wn, rel=1e-12)
        eps = 0.5 * (F - np.eye(3) + (F - np.eye(3)).T)
        assert oracle.law_W("linear", lam, mu, g) == pytest.approx(0.5 * lam * np.trace(eps) ** 2 + mu * np.sum(eps * eps),
                                                                   rel=1e-13)


@pytest.mark.parametrize("law", ["neohookean", "linear"])
def test_factorised_tangent_form(oracle, law):
    """The factorised tangent storage rests on T[(J,c),(L,k)] = mu d_ck d_JL + m A_Jk A_Lc + lam A_Jc A_Lk with
    A = F^-1 and m = mu - lam ln J (neo-Hookean) or A = I and m = mu (linear); and, in reference coordinates with
    H = iJt' A and g = iJt' iJt, D = dV iJt' T iJt = dV (mu g_KM d_ck + m H_Kk H_Mc + lam H_Kc H_Mk)."""
    rng = np.random.default_rng(6)
    lam, mu = 2.0, 1.3
    for _ in range(4):
        g = _rand_g(rng, 0.2)
        F = _F(g)
        A = np.linalg.inv(F) if law == "neohookean" else np.eye(3)
        m = mu - lam * np.log(np.linalg.det(F)) if law == "neohookean" else mu
        d = np.eye(3)
        T = oracle.law_T(law, lam, mu, g).reshape(9, 9, order="F")  # T[J + 3c, L + 3k]
        Tf = np.zeros((9, 9))
        for c in range(3):
            for J in range(3):
                for k in range(3):
                    for L in range(3):
                        Tf[J + 3 * c, L + 3 * k] = (mu * d[c, k] * d[J, L] + m * A[J, k] * A[L, c]
                                                    + lam * A[J, c] * A[L, k])
        assert np.allclose(Tf, T, atol=1e-13, rtol=1e-13)
        iJt = np.eye(3) + 0.3 * rng.standard_normal((3, 3))  # iJt[J, K]
        dV = 0.37
        D = np.zeros((9, 9))
        for c in range(3):
            for K in range(3):
                for k in range(3):
                    for M in range(3):
                        D[K + 3 * c, M + 3 * k] = dV * sum(iJt[J, K] * T[J + 3 * c, L + 3 * k] * iJt[L, M]
                                                           for J in range(3) for L in range(3))
        H = iJt.T @ A
        G = iJt.T @ iJt
        Df = np.zeros((9, 9))
        for c in range(3):
            for K in range(3):
                for k in range(3):
                    for M in range(3):
                        Df[K + 3 * c, M + 3 * k] = dV * (mu * G[K, M] * d[c, k] + m * H[K, k] * H[M, c]
                                                         + lam * H[K, c] * H[M, k])
        assert np.allclose(Df, D, atol=1e-13, rtol=1e-12)
