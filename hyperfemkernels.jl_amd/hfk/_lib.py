"""ctypes binding of libhfk.so (the C ABI of include/hfk.h).

The library is built in-tree by ``hfk.build``; loading fails loudly when it is missing -- there is no CPU
fallback for any operator of this package.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhfk.so")

HFK_OK = 0
STATUS = {-1: "ArgumentError", -2: "DimensionMismatch", -3: "ErrorException", -4: "HIP error", -5: "allocation failure"}

STORE = {"full64": 0, "sym64": 1, "full32": 2, "sym32": 3, "fac64": 4, "fac32": 5, "afac64": 6, "afac32": 7}
LAWS = {"svk": 0, "neohookean": 1, "linear": 2}

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)
_vp = C.c_void_p

# name -> (restype, argtypes); every symbol that include/hfk.h declares
SIGNATURES = {
    "hfk_version": (C.c_int, []),
    "hfk_open": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "hfk_close": (C.c_int, [_vp]),
    "hfk_last_error": (C.c_char_p, [_vp]),
    "hfk_synchronize": (C.c_int, [_vp]),
    "hfk_op_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int64, _i32p, _i32p, _i32p, _f64p, C.c_int64, C.c_int64,
                                C.c_int, C.c_double, C.c_double, C.POINTER(_vp)]),
    "hfk_op_destroy": (C.c_int, [_vp]),
    "hfk_op_set_dirichlet": (C.c_int, [_vp, _f64p]),
    "hfk_op_set_refinement": (C.c_int, [_vp, _vp, _i32p, _i32p]),
    "hfk_op_set_parameters": (C.c_int, [_vp, C.c_double, C.c_double]),
    "hfk_op_set_mu_field": (C.c_int, [_vp, _f64p]),
    "hfk_op_set_traction": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f64p]),
    "hfk_op_set_load_factor": (C.c_int, [_vp, C.c_double]),
    "hfk_op_set_robin": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, _f64p]),
    "hfk_op_set_skeleton": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, _i32p, _i32p, _i32p, C.c_char_p, C.c_char_p,
                                      _f64p, C.c_int]),
    "hfk_op_skeleton_colors": (C.c_int, [_vp, _i32p]),
    "hfk_skel_form_check": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64]),
    "hfk_op_set_boundary_form": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, C.c_char_p, C.c_char_p, _f64p, C.c_int]),
    "hfk_op_traction_load": (C.c_int, [_vp, _f64p]),
    "hfk_op_set_mu_function": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _f64p, C.c_int64]),
    "hfk_op_mu_field": (C.c_int, [_vp, _f64p]),
    "hfk_energy": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_energy_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_energy_noalloc_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_pointfn_create": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "hfk_pointfn_destroy": (C.c_int, [_vp]),
    "hfk_extremum": (C.c_int, [_vp, _vp, _f64p, _f64p, _f64p, _f64p, C.c_int64, C.c_int, _f64p]),
    "hfk_op_info": (C.c_int, [_vp, _i64p]),
    "hfk_op_colors": (C.c_int, [_vp, _i32p]),
    "hfk_op_set_deterministic": (C.c_int, [_vp, C.c_int]),
    "hfk_law_create": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.POINTER(_vp)]),
    "hfk_law_destroy": (C.c_int, [_vp]),
    "hfk_law_check": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64]),
    "hfk_law_info": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    "hfk_op_set_law": (C.c_int, [_vp, _vp]),
    "hfk_residual": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_residual_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_mf_prepare": (C.c_int, [_vp, _f64p, C.c_int]),
    "hfk_mf_prepare_dev": (C.c_int, [_vp, _vp, C.c_int, _vp]),
    "hfk_mf_apply": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_mf_apply_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_mf_tangents": (C.c_int, [_vp, _f64p]),
    "hfk_pattern_nnz": (C.c_int, [_vp, _i64p]),
    "hfk_pattern": (C.c_int, [_vp, _i64p, _i64p, _i32p]),
    "hfk_jacobian": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_jacobian_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_block_jacobi_update": (C.c_int, [_vp, _i32p]),
    "hfk_block_jacobi_apply": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_block_jacobi_apply_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_block_jacobi_blocks": (C.c_int, [_vp, _i64p, _i32p, _f64p]),
    "hfk_cheb_create": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, C.POINTER(_vp)]),
    "hfk_cheb_destroy": (C.c_int, [_vp]),
    "hfk_cheb_update": (C.c_int, [_vp, _i32p, _f64p]),
    "hfk_cheb_smooth": (C.c_int, [_vp, _f64p, _f64p, C.c_int, _f64p]),
    "hfk_cheb_smooth_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, _vp]),
    "hfk_eigenvalue_bounds": (C.c_int, [_vp, C.c_int, _f64p, _f64p]),
    "hfk_transfer_create": (C.c_int, [_vp, _vp, C.POINTER(_vp)]),
    "hfk_transfer_destroy": (C.c_int, [_vp]),
    "hfk_prolongate": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_restrict": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_inject": (C.c_int, [_vp, _f64p, _f64p, _f64p, _f64p]),
    "hfk_pmg_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                 C.c_int, C.c_double, C.c_int, C.POINTER(_vp)]),
    "hfk_pmg_destroy": (C.c_int, [_vp]),
    "hfk_pmg_update": (C.c_int, [_vp, _i32p]),
    "hfk_pmg_set_coarse_every": (C.c_int, [_vp, C.c_int]),
    "hfk_pmg_coarse_builds": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "hfk_pmg_set_coarse_precision": (C.c_int, [_vp, C.c_int]),
    "hfk_pmg_set_fallback": (C.c_int, [_vp, C.c_int, C.c_double]),
    "hfk_pmg_fallbacks": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "hfk_pmg_vcycle": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_pmg_vcycle_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_pmg_info": (C.c_int, [_vp, _f64p]),
    "hfk_pmg_coarse_matvec": (C.c_int, [_vp, _f64p, _f64p, C.c_int]),
    "hfk_net_create": (C.c_int, [_vp, C.c_int, C.c_int, _i32p, C.c_int64, C.c_double, C.c_uint64, C.POINTER(_vp)]),
    "hfk_net_destroy": (C.c_int, [_vp]),
    "hfk_net_num_params": (C.c_int64, [_vp]),
    "hfk_net_get_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "hfk_net_set_weights": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "hfk_net_apply": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_net_apply_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_net_gradient": (C.c_int, [_vp, _f64p, _f64p, C.c_int, C.POINTER(C.c_float), _f64p]),
    "hfk_net_train_dev": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _f64p, _vp]),
    "hfk_pmg_set_coarse_network": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int]),
    "hfk_pmg_coarse_pairs_dev": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.c_int, C.c_double, C.c_int, _vp]),
    "hfk_net_set_aux": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "hfk_pmg_train_coarse_network": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _f64p,
                                               _vp]),
    "hfk_pmg_coarse_network_error": (C.c_int, [_vp, _vp, _vp, C.c_int, _f64p, _vp]),
    "hfk_pmg_coarse_network_energy_gradient": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(C.c_float), _f64p, _vp]),
    "hfk_oseen_create": (C.c_int, [_vp, C.c_int64, _i32p, _i32p, _i32p, _i32p, _f64p, C.c_int, _f64p, _f64p, C.c_int64,
                                   C.c_int64, C.c_double, C.c_double, C.c_double, C.POINTER(_vp)]),
    "hfk_oseen_destroy": (C.c_int, [_vp]),
    "hfk_oseen_set_form": (C.c_int, [_vp, C.c_char_p, C.c_char_p, _f64p, C.c_int]),
    "hfk_oseen_set_blocks": (C.c_int, [_vp, _i32p, _i32p]),
    "hfk_oseen_form_check": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64]),
    "hfk_oseen_set_fields": (C.c_int, [_vp, _f64p]),
    "hfk_oseen_set_dirichlet": (C.c_int, [_vp, _f64p]),
    "hfk_oseen_set_laplacian": (C.c_int, [_vp, _f64p]),
    "hfk_oseen_step_fields": (C.c_int, [_vp, _f64p, _f64p, _f64p, _f64p, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int]),
    "hfk_oseen_step_fields_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_int, _vp]),
    "hfk_oseen_set_outlet": (C.c_int, [_vp, C.c_int64, _i32p, _f64p, C.c_int, _f64p, C.c_int, _f64p, C.c_double,
                                       C.c_double]),
    "hfk_oseen_set_outlet_pressure": (C.c_int, [_vp, C.c_double]),
    "hfk_oseen_set_parameters": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double]),
    "hfk_oseen_set_deterministic": (C.c_int, [_vp, C.c_int]),
    "hfk_oseen_info": (C.c_int, [_vp, _i64p]),
    "hfk_oseen_colors": (C.c_int, [_vp, _i32p]),
    "hfk_oseen_pattern_nnz": (C.c_int, [_vp, _i64p]),
    "hfk_oseen_pattern": (C.c_int, [_vp, _i64p, _i64p, _i32p]),
    "hfk_oseen_residual": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_oseen_residual_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_oseen_jacobian": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_oseen_jacobian_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_oseen_residual_jacobian_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "hfk_krylov_solve": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _f64p, _f64p, C.c_double, C.c_double, C.c_int, C.c_int,
                                   _i32p, _f64p]),
    "hfk_krylov_solve_dev": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, C.c_double, C.c_double, C.c_int, C.c_int,
                                       _i32p, _f64p, _vp]),
    "hfk_sys_create": (C.c_int, [_vp, C.c_int64, _i64p, _i64p, C.POINTER(_vp)]),
    "hfk_sys_destroy": (C.c_int, [_vp]),
    "hfk_sys_set_blocks": (C.c_int, [_vp, C.c_int64, _i32p]),
    "hfk_sys_update_dev": (C.c_int, [_vp, _vp, _i64p, _vp]),
    "hfk_sys_matvec_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_sys_blocks_apply_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_sys_smooth_dev": (C.c_int, [_vp, C.c_int, C.c_double, _vp, _vp, _vp]),
    "hfk_sys_blocks": (C.c_int, [_vp, _f64p, _f64p]),
    "hfk_sys_set_patches": (C.c_int, [_vp, C.c_int64, _i64p, _i32p, C.POINTER(C.c_int8)]),
    "hfk_sys_patches_apply_dev": (C.c_int, [_vp, _vp, _vp, _vp]),
    "hfk_sys_patch_inverses": (C.c_int, [_vp, _f64p]),
    "hfk_sys_solve_dev": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, _vp, _vp, C.c_double, C.c_double,
                                    C.c_int, C.c_int, _i32p, _f64p, _vp]),
}

_lib = None


class HFKError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, 'error')} ({status}): {msg}")
        self.status = status


def load(path: str = LIB_PATH):
    """Load libhfk.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"libhfk.so not found at {path}: build it with hfk.build.build() "
                          "(there is no CPU fallback)")
    # hiprtc of the HIP runtime already in the process: torch bundles both (same soname as /opt/rocm's)
    torch = sys.modules.get("torch")
    if torch is not None:
        rtc = os.path.join(os.path.dirname(torch.__file__), "lib", "libhiprtc.so")
        if os.path.exists(rtc):
            C.CDLL(rtc, mode=C.RTLD_GLOBAL)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(status: int, ctx=None):
    if status != HFK_OK:
        lib = load()
        msg = lib.hfk_last_error(ctx)
        raise HFKError(status, msg.decode() if msg else "")


def ptr(a, ctype):
    """Pointer to the data of a contiguous numpy array."""
    return a.ctypes.data_as(C.POINTER(ctype))
