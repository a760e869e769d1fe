"""CPU checks of the C ABI: libhfk.so loads, exports every function include/hfk.h declares with the ctypes
signatures the host mirror binds, and reports errors (no GPU work is attempted here)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "hfk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hfk_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    from hfk import _lib

    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "ctypes signatures out of sync with include/hfk.h"
    assert lib.hfk_version() >= 1


def test_julia_binding_calls_declared_symbols():
    """Every entry point the Julia ccall layer binds is declared in include/hfk.h (and so exported, above)."""
    src = open(os.path.join(ROOT, "hyperfemkernels.jl_amd", "julia", "HyperFEMKernelsMI355X.jl")).read()
    called = set(re.findall(r":(hfk_[a-z0-9_]+)\b", src))
    assert len(called) >= 20
    missing = sorted(called - set(declared()))
    assert not missing, missing


def test_library_is_gfx950_only():
    from hfk import _lib

    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_error_paths_without_device():
    import torch

    from hfk import _lib

    lib = _lib.load()
    if torch.cuda.is_available():
        pytest.skip("a device is present: covered by the GPU tests")
    h = C.c_void_p()
    st = lib.hfk_open(0, C.byref(h))
    assert st != 0 and not h.value
    assert b"device" in lib.hfk_last_error(None)
    assert lib.hfk_op_create(None, 3, 3, 1, None, None, None, None, 1, 0, 0, 1.0, 1.0, C.byref(h)) == -1
