"""The face kernels compile with hiprtc for gfx950 on the host (no device): the built-in gradient-jump penalty
(skeleton) and follower pressure (boundary) for every supported (Q, P), nonlinear forms given as HIP source, and a
broken form reported with the compiler's log; the C-ABI validates (Q, P), the number of sides and the form's name."""
import ctypes as C

import pytest

from skeleton_forms import NONLINEAR_BOUNDARY_SOURCE, NONLINEAR_FACET_SOURCE


@pytest.fixture(scope="module")
def lib():
    from hfk._lib import load

    return load()


def check(lib, Q, P, source=None, name=None, sides=2):
    log = C.create_string_buffer(4096)
    rc = lib.hfk_skel_form_check(Q, P, sides, source.encode() if source else None, name.encode() if name else None,
                                 log, 4096)
    return rc, log.value.decode()


@pytest.mark.parametrize("Q,P", [(2, 2), (2, 3), (3, 3), (3, 4), (4, 4)])
@pytest.mark.parametrize("sides", [1, 2])
def test_builtin_compiles(lib, Q, P, sides):
    assert check(lib, Q, P, sides=sides)[0] == 0


def test_compiled_form_and_errors(lib):
    assert check(lib, 3, 3, NONLINEAR_FACET_SOURCE, "NonlinearFacet")[0] == 0
    assert check(lib, 3, 3, NONLINEAR_BOUNDARY_SOURCE, "NonlinearBoundary", sides=1)[0] == 0
    rc, log = check(lib, 3, 3, NONLINEAR_FACET_SOURCE.replace("parts(ds, dJ, dum);", "parts(ds, dJ, dum)"), "NonlinearFacet")
    assert rc != 0 and "error" in log
    assert check(lib, 3, 3, NONLINEAR_FACET_SOURCE, "not an identifier")[0] != 0
    assert check(lib, 5, 5)[0] != 0
    assert check(lib, 3, 3, sides=3)[0] != 0
