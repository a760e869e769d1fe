"""CPU checks of the law code generator (hfk.codegen, the analogue of src/codegen.jl:122-198): the stress and
tangent derived symbolically from the SVK and neo-Hookean energies equal the oracle's hand-written laws (which are
pinned by closed forms in test_oracle_pointwise.py); a stress-defined law gets the same tangent as its energy;
energy-derived tangents are symmetric; a law the library does not build in (compressible Mooney-Rivlin) matches
central differences of its energy and stress; the emitted HIP source compiles with hiprtc (no device)."""
import numpy as np
import pytest

from hfk import codegen as cg
from hfk.laws import check_law


def oracle_layout(F):
    """g[J + 3c] for F[c, J] = delta + g."""
    return (F - np.eye(3)).reshape(-1)


def rand_F(rng, scale=0.2):
    return np.eye(3) + scale * rng.standard_normal((3, 3))


def P_from_oracle(oracle, law, lam, mu, F):
    return oracle.law_P(law, lam, mu, oracle_layout(F)).reshape(3, 3)  # R[J + 3c] = P[c, J]


def T_from_oracle(oracle, law, lam, mu, F):
    t = oracle.law_T(law, lam, mu, oracle_layout(F))
    T = np.zeros((3, 3, 3, 3))
    for c in range(3):
        for J in range(3):
            for k in range(3):
                for L in range(3):
                    T[c, J, k, L] = t[(J + 3 * c) + 9 * (L + 3 * k)]
    return T


@pytest.mark.parametrize("law,energy", [("svk", cg.svk_energy), ("neohookean", cg.neohooke_energy)])
def test_energy_laws_match_oracle(oracle, law, energy):
    pl = cg.from_energy(energy)
    rng = np.random.default_rng(7)
    for _ in range(4):
        F = rand_F(rng)
        P, T = pl.evaluate(F, 2.0, 1.3)
        Pr, Tr = P_from_oracle(oracle, law, 2.0, 1.3, F), T_from_oracle(oracle, law, 2.0, 1.3, F)
        assert np.abs(P - Pr).max() <= 1e-13 * np.abs(Pr).max()
        assert np.abs(T - Tr).max() <= 1e-13 * np.abs(Tr).max()
        assert np.allclose(T, T.transpose(2, 3, 0, 1), atol=1e-13)  # hyperelastic: symmetric tangent
        W = pl.energy(F, 2.0, 1.3)
        assert W == pytest.approx(oracle.law_W(law, 2.0, 1.3, oracle_layout(F)), rel=1e-13)


def test_stress_defined_law_gets_the_same_tangent():
    def svk_stress(F, lam, mu):
        E = (F.T() @ F - cg.Mat3.identity()) * 0.5
        S = cg.Mat3.identity() * (lam * cg.tr(E)) + E * (2.0 * mu)
        return F @ S

    a, b = cg.from_stress(svk_stress), cg.from_energy(cg.svk_energy)
    F = rand_F(np.random.default_rng(3))
    Pa, Ta = a.evaluate(F, 2.0, 1.0)
    Pb, Tb = b.evaluate(F, 2.0, 1.0)
    assert np.allclose(Pa, Pb, atol=1e-14) and np.allclose(Ta, Tb, atol=1e-13)


def mooney_rivlin(F, lam, mu):
    """W = mu/2 (I1 - 3) + mu/4 (I2 - 3) - 2 mu lnJ + lam/2 (lnJ)^2 (stress free at F = I)."""
    C = F.T() @ F
    I1 = cg.tr(C)
    I2 = 0.5 * (I1 * I1 - cg.tr(C @ C))
    lnJ = cg.log(cg.det(F))
    return mu * 0.5 * (I1 - 3.0) + mu * 0.25 * (I2 - 3.0) - 2.0 * mu * lnJ + lam * 0.5 * lnJ * lnJ


def energy_value(F, lam, mu):
    C = F.T @ F
    I1, I2 = np.trace(C), 0.5 * (np.trace(C) ** 2 - np.trace(C @ C))
    lnJ = np.log(np.linalg.det(F))
    return mu * 0.5 * (I1 - 3) + mu * 0.25 * (I2 - 3) - 2 * mu * lnJ + lam * 0.5 * lnJ ** 2


def test_extra_law_against_finite_differences():
    pl = cg.from_energy(mooney_rivlin)
    rng = np.random.default_rng(11)
    lam, mu, h = 2.0, 0.7, 1e-6
    for _ in range(3):
        F = rand_F(rng)
        P, T = pl.evaluate(F, lam, mu)
        for c in range(3):
            for J in range(3):
                E = np.zeros((3, 3))
                E[c, J] = h
                fd = (energy_value(F + E, lam, mu) - energy_value(F - E, lam, mu)) / (2 * h)
                assert fd == pytest.approx(P[c, J], rel=1e-7, abs=1e-9)
                Pp, _ = pl.evaluate(F + E, lam, mu)
                Pm, _ = pl.evaluate(F - E, lam, mu)
                assert np.allclose((Pp - Pm) / (2 * h), T[:, :, c, J], rtol=1e-6, atol=1e-8)
    P0, T0 = pl.evaluate(np.eye(3), lam, mu)
    assert np.abs(P0).max() < 1e-15  # stress free reference configuration


def test_generated_source_compiles():
    for name, energy in (("GenSVK", cg.svk_energy), ("GenNeoHooke", cg.neohooke_energy), ("GenMooney", mooney_rivlin)):
        src = cg.from_energy(energy).source(name)
        assert f"struct {name}" in src and "stress" in src and "tangent" in src and "energy()" in src
        check_law(src, name)


def test_extra_law_linearises_to_lame():
    """At F = I the Mooney-Rivlin energy above linearises to the isotropic linear law with shear modulus 3 mu / 2
    and first Lame parameter lam + mu (worked by hand from W's derivatives at the identity)."""
    _, T = cg.from_energy(mooney_rivlin).evaluate(np.eye(3), 2.0, 0.8)
    d = np.eye(3)
    lam, G = 2.8, 1.2
    ref = lam * np.einsum("cJ,kL->cJkL", d, d) + G * (np.einsum("ck,JL->cJkL", d, d) + np.einsum("cL,Jk->cJkL", d, d))
    assert np.abs(T - ref).max() < 1e-14
