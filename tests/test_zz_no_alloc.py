"""GPU test of hfk_energy_noalloc_dev, the energy entry point that allocates nothing (its partial sums live in the
library's static per-device scratch), called through the C ABI with device pointers.

It must give hfk_energy_dev's value bit for bit (same kernels, same fixed summation order) and the oracle's energy
to 1e-12:
* two calls in a row, the second on a smaller operator (fewer partial sums than the first left in the scratch);
* with a traction load and Robin faces (their partial sums follow the batches' in the scratch);
* a compiled law is refused (HFK_ESTATE) until the operator has loaded its kernels, then agrees too.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from hfk import HyperelasticForm, KernelOperator
from hfk._lib import HFK_OK
from hfk.functionals import KernelFunctional
from hfk.mesh import boundary_faces, dirichlet_closed_x0, hex_space

pytestmark = pytest.mark.gpu


def make(hip_ctx, oracle, n, seed, faces):
    sp = hex_space(n, 2, dirichlet=dirichlet_closed_x0, perturb=0.1, seed=seed)
    rng = np.random.default_rng(seed)
    dv = 0.01 * rng.standard_normal(sp.ndir)
    x = 0.01 * rng.standard_normal(sp.nfree)
    op = KernelOperator(HyperelasticForm("neohookean", 2.0, 1.0, degree=4), sp, jacobian="matrixfree",
                        dirichlet_values=dv, ctx=hip_ctx)
    orc = oracle.OracleOp(sp, "neohookean", 2.0, 1.0, dvals=dv)
    ref = None
    if faces:
        fc1, fl1 = boundary_faces(sp, 0, 1)
        fc2, fl2 = boundary_faces(sp, 2, 0)
        op.set_traction(fc1, fl1, [0.3, -0.2, 0.7])
        op.set_load_factor(1.3)
        op.set_robin(fc2, fl2, 0.8)
        ff, fd = orc.face_load(fc1, fl1, np.tile([0.3, -0.2, 0.7], (fc1.size, 1)))
        M = orc.face_mass(fc2, fl2, np.full(fc2.size, 0.8))
        u = np.concatenate([x, dv])
        ref = orc.energy(x) - 1.3 * (ff @ x + fd @ dv) + 0.5 * u @ M @ u
    else:
        ref = orc.energy(x)
    return sp, x, op, ref


def energies(hip_ctx, op, x):
    """(hfk_energy_dev, hfk_energy_noalloc_dev) of op at x, both with device pointers on the context's stream."""
    lib = op.lib
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert lib.hfk_energy_dev(op.handle, C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr()), None) == HFK_OK
    rc = lib.hfk_energy_noalloc_dev(op.handle, C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr() + 8), None)
    assert rc == HFK_OK, lib.hfk_last_error(hip_ctx.handle)
    hip_ctx.synchronize()
    e = out.cpu().numpy()
    return float(e[0]), float(e[1])


def test_noalloc_energy_equals_energy_dev(hip_ctx, oracle):
    print("noalloc energy: 6x4x3 Q2 with traction and Robin faces, then 2x2x1 without", flush=True)
    sp, x, op, ref = make(hip_ctx, oracle, (6, 4, 3), 5, True)
    e_dev, e_na = energies(hip_ctx, op, x)
    assert e_na == e_dev
    assert e_na == pytest.approx(ref, rel=1e-12)
    sp2, x2, op2, ref2 = make(hip_ctx, oracle, (2, 2, 1), 6, False)
    e_dev2, e_na2 = energies(hip_ctx, op2, x2)
    assert e_na2 == e_dev2
    assert e_na2 == pytest.approx(ref2, rel=1e-12)
    # the first operator again, after the smaller one: the scratch keeps no state between calls
    assert energies(hip_ctx, op, x) == (e_dev, e_na)


def test_noalloc_energy_compiled_law(hip_ctx, oracle):
    from hfk.laws import SOURCES, CompiledLaw

    print("noalloc energy: compiled neo-Hookean law", flush=True)
    sp, x, op, ref = make(hip_ctx, oracle, (3, 2, 2), 7, False)
    src, name = SOURCES["neohookean"]
    op.set_law(CompiledLaw(src, name, ctx=hip_ctx))
    lib = op.lib
    xd = torch.tensor(x, dtype=torch.float64, device="cuda")
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = lib.hfk_energy_noalloc_dev(op.handle, C.c_void_p(xd.data_ptr()), C.c_void_p(out.data_ptr()), None)
    assert rc == -3  # HFK_ESTATE: the law's kernels are not loaded yet
    KernelFunctional(op)(x)  # loads them
    e_dev, e_na = energies(hip_ctx, op, x)
    assert e_na == e_dev
    assert e_na == pytest.approx(ref, rel=1e-12)
