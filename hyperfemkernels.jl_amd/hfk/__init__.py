"""hfk — host mirror of HyperFEMKernels.jl's matrix-free and assembly hot path over libhfk (HIP, gfx950).

Modules: ``mesh`` (Cartesian hex Q_k spaces, the stand-in for Gridap's spaces), ``synthetic`` (SplitMix64
inputs), ``operator`` (KernelOperator, MatrixFreeJacobian), ``laws``
(pointwise laws compiled at run time with hiprtc), ``codegen`` (laws generated from an expression tree), ``build`` (hipcc build of libhfk.so), ``_lib``
(ctypes binding).  Importing this package does not load the library; the first operator does, and fails loudly
if it has not been built.
"""
from .mesh import HexSpace, hex_space, table_node_order  # noqa: F401
from .laws import CompiledLaw, check_law  # noqa: F401
from .operator import Context, HyperelasticForm, KernelOperator, MatrixFreeJacobian  # noqa: F401

__version__ = "0.1.0"
