"""Misuse is rejected on the host before any device work, as the reference throws ArgumentError for it
(test/solvers/test_matrixfree_solver.jl:58-60, test_chebyshev.jl:124-125, test_pmultigrid.jl:152): invalid
Krylov method, tolerance or iteration limit, Chebyshev degree or range, a one-level hierarchy, an unknown Jacobian
kind or coarse-network mode."""
import pytest

from hfk.learned import MODES
from hfk.solvers import Chebyshev, MatrixFreeSolver, PMultigrid


def test_matrixfree_solver_arguments():
    with pytest.raises(ValueError):
        MatrixFreeSolver(krylov="bicgstab")
    with pytest.raises(ValueError):
        MatrixFreeSolver(rtol=1.0)
    with pytest.raises(ValueError):
        MatrixFreeSolver(maxiter=0)
    with pytest.raises(ValueError):
        MatrixFreeSolver(restart=0)
    s = MatrixFreeSolver()
    assert s.krylov == "gmres" and s.restart == 50 and s.maxiter == 500  # the reference's defaults
    with pytest.raises(RuntimeError, match="MatrixFreeJacobian"):
        s.solve(None, object(), None)


def test_chebyshev_and_pmultigrid_arguments():
    with pytest.raises(ValueError):
        Chebyshev(degree=0)
    with pytest.raises(ValueError):
        Chebyshev(range=1.0)
    with pytest.raises(ValueError):
        PMultigrid([object()])
    assert Chebyshev(degree=3).damping() > 0


def test_coarse_network_modes():
    assert MODES == {"off": 0, "warmstart": 1, "replace": 2, "precondition": 3, "cycle": 4}
