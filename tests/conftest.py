import os
import sys

import pytest
# torch ships its own HIP runtime; load it before libhfk (which links the system ROCm runtime) so that tests which
# use torch device tensors as plumbing find the device (the bench imports torch first for the same reason)
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hyperfemkernels.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (run on the MI355X box)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as orc

    orc.load()
    return orc


@pytest.fixture(scope="session")
def hip_ctx():
    import hfk

    return hfk.Context(0)
