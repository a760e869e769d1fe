"""Host-side mesh helpers (no GPU): quadrature-point coordinates used for heterogeneous shear-modulus fields."""
import numpy as np

from hfk.mesh import dirichlet_closed_x0, hex_space
from oracle.oracle import gauss01


def test_quadrature_points_cartesian():
    sp = hex_space((3, 2, 4), 2, domain=(0.0, 1.5, -1.0, 1.0, 0.0, 2.0), dirichlet=dirichlet_closed_x0)
    for P in (2, 3, 4):
        g, _ = gauss01(P)
        pts = np.array([[g[a], g[b], g[c]] for c in range(P) for b in range(P) for a in range(P)])
        lo, hi = sp.coords[:, 0, :], sp.coords[:, 7, :]
        ref = lo[:, None, :] + pts[None, :, :] * (hi - lo)[:, None, :]
        np.testing.assert_allclose(sp.quadrature_points(P), ref, rtol=0, atol=1e-14)


def test_quadrature_points_perturbed_cells():
    sp = hex_space((3, 3, 3), 1, perturb=0.2, seed=5)
    xq = sp.quadrature_points(3)
    # the centre point of each cell is the mean of its 8 vertices (trilinear map at xi = 1/2)
    np.testing.assert_allclose(xq[:, 13, :], sp.coords.mean(axis=1), rtol=0, atol=1e-14)
    # every point lies inside the bounding box of its cell
    assert np.all(xq >= sp.coords.min(axis=1)[:, None, :] - 1e-14)
    assert np.all(xq <= sp.coords.max(axis=1)[:, None, :] + 1e-14)
