"""Point forms generated from a captured velocity-pressure integrand (hfk.formgen), checked on the host.

* The Oseen SUPG/PSPG/LSIC integrand of test/common.jl:207-220 written with the DAG: on random affine cells and
  point data, the element matrices built from its tangent -- both the derived 16 x 16 slot tangent and the 50
  projected factors the device kernel contracts -- equal the oracle's element matrices (term-by-term tensor
  algebra, 1e-12), and the element residuals from its test-slot coefficients equal K x_e - load of the oracle.
* A stabilisation variant the library does not build in (the transient SUPG parameter with c0^2 added under the
  root, scaled by a parameter of the form) equals the oracle with that tau written in numpy.
* Forms outside the factorised class (an anisotropic component coupling) and forms that are not bilinear are
  refused; the emitted source compiles with hiprtc (no device).
"""
import numpy as np
import pytest

from hfk import formgen as fg


def cells_and_data(seed, nc=3):
    from oracle import oseen as O  # noqa: F401

    rng = np.random.default_rng(seed)
    ref = np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    coords = np.array([ref @ (np.eye(3) + 0.3 * rng.standard_normal((3, 3))).T + rng.standard_normal(3)
                       for _ in range(nc)])
    from hfk.tetmesh import tet_quadrature_deg5

    qp, qw = tet_quadrature_deg5()
    fields = rng.standard_normal((nc, len(qw), 9))
    lap = rng.standard_normal((nc, len(qw), 3))
    return rng, coords, qp, qw, fields, lap


def element_from_form(form, coords, qp, qw, fields, lap, prm, state_e=None, projected=False):
    """Element matrices sum_q dV phi_a' T_q phi_b and residual vectors sum_q dV phi_a . co_q from the form."""
    from oracle import oseen as O

    Nr, Gr = O.basis(qp)
    nc = coords.shape[0]
    K = np.zeros((nc, 34, 34))
    R = np.zeros((nc, 34))
    for c in range(nc):
        J = (coords[c, 1:] - coords[c, :1]).T
        det = np.linalg.det(J)
        Jit = np.linalg.inv(J).T
        h = (6 * np.sqrt(2) * abs(det) / 6.0) ** (1.0 / 3.0)
        for q in range(len(qw)):
            gx = Gr[q] @ Jit.T  # (14, 3) physical gradients
            phi = np.zeros((34, 16))
            for i in range(3):
                for a in range(10):
                    phi[i * 10 + a, i] = Nr[q, a]
                    phi[i * 10 + a, 3 + 3 * i:6 + 3 * i] = gx[a]
            for v in range(4):
                phi[30 + v, 12] = Nr[q, 10 + v]
                phi[30 + v, 13:16] = gx[10 + v]
            st = np.zeros(16) if state_e is None else phi.T @ state_e[c]
            d = np.concatenate([fields[c, q], lap[c, q]])
            co, T, k = form.evaluate(st, d, h, prm)
            if projected:
                T = form.reconstruct(k)
            dV = abs(det) * qw[q]
            K[c] += dV * phi @ T @ phi.T
            R[c] += dV * phi @ co
    return K, R


@pytest.mark.parametrize("projected", [False, True])
def test_oseen_form_equals_oracle(oracle, projected):
    from oracle import oseen as O

    rng, coords, qp, qw, fields, lap = cells_and_data(1)
    rho, mu, c0 = 1.3, 0.05, 150.0
    form = fg.oseen_supg()
    K, _ = element_from_form(form, coords, qp, qw, fields, lap, [rho, mu, c0], projected=projected)
    Ko, load = O.element_system(coords, qp, qw, fields, rho, mu, c0, lap=lap)
    assert np.abs(K - Ko).max() <= 1e-12 * np.abs(Ko).max()
    xe = rng.standard_normal((coords.shape[0], 34))
    _, R = element_from_form(form, coords, qp, qw, fields, lap, [rho, mu, c0], state_e=xe)
    Ro = np.einsum("cab,cb->ca", Ko, xe) - load
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()


def tau_transient(D):
    """SUPG with the time scale under the root: ts = ((ct c0)^2 + (2|un|/h)^2 + 9 (4 mu / (rho h^2))^2)^(-1/2)."""
    unn = fg.norm(D.un)
    r0 = D.param("ct") * D.c0
    r1 = 2.0 * unn / D.h
    r2 = 3.0 * (4.0 * D.mu / (D.rho * D.h * D.h))
    ts = 1.0 / fg.sqrt(r0 * r0 + r1 * r1 + r2 * r2)
    return ts, ts / D.rho, D.rho * unn * D.h * 0.5


def test_tau_variant_equals_oracle_with_that_tau(oracle):
    from oracle import oseen as O

    rng, coords, qp, qw, fields, lap = cells_and_data(2)
    rho, mu, c0, ct = 1.0, 0.01, 100.0, 0.02
    form = fg.oseen_supg(tau=tau_transient, params=("ct",), name="OseenTransientSupg")

    def tau_np(unn, h, rho, mu, c0):
        ts = ((ct * c0) ** 2 + (2 * unn / h) ** 2 + 9 * (4 * mu / (rho * h ** 2)) ** 2) ** -0.5
        return ts, ts / rho, rho * unn * h * 0.5

    K, _ = element_from_form(form, coords, qp, qw, fields, lap, [rho, mu, c0, ct], projected=True)
    Ko, _ = O.element_system(coords, qp, qw, fields, rho, mu, c0, lap=lap, tau=tau_np)
    assert np.abs(K - Ko).max() <= 1e-12 * np.abs(Ko).max()
    Kr, _ = O.element_system(coords, qp, qw, fields, rho, mu, c0, lap=lap)
    assert np.abs(K - Kr).max() > 1e-6 * np.abs(Ko).max()  # the variant differs from the built-in integrand


def test_forms_outside_the_class_are_refused():
    def aniso(U, V, D):  # couples velocity component 0 of the test with component 1 of the trial
        return fg.dot(U.u, V.u) + U.u[1] * V.u[0]

    with pytest.raises(ValueError, match="factorised"):
        fg.OseenPointForm(aniso, lambda V, D: fg.const(0.0), name="Aniso")

    def nonlinear(U, V, D):
        return fg.dot(U.u, U.u) * fg.dot(D.w, V.u)

    with pytest.raises(ValueError, match="bilinear"):
        fg.OseenPointForm(nonlinear, lambda V, D: fg.const(0.0), name="Nonlin")


def test_form_source_compiles():
    fg.check_compiles(fg.oseen_supg())
    fg.check_compiles(fg.oseen_supg(tau=tau_transient, params=("ct",), name="OseenTransientSupg"))
    bad = fg.oseen_supg(name="Broken")
    bad_src = bad.source().replace("co[0] =", "co[0] = undefined_symbol +", 1)
    import ctypes as C

    from hfk import _lib

    log = C.create_string_buffer(1 << 14)
    assert _lib.load().hfk_oseen_form_check(bad_src.encode(), b"Broken", log, len(log)) == -1
    assert b"undefined_symbol" in log.value
