"""ctypes binding of libnlfh.so (the C ABI in include/nlfh.h).

The library is the only compute path: if it is missing or fails to load, every call raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnlfh.so")

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# (name, restype, argtypes): every symbol of include/nlfh.h
SIGNATURES = [
    ("nlfh_dmap_create", ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.POINTER(_vp)]),
    ("nlfh_dmap_destroy", ctypes.c_int, [_vp]),
    ("nlfh_dmap_smooth", ctypes.c_int, [_vp, _vp, _vp, _vp]),
    ("nlfh_dmap_rho", ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, _c_double_p, _vp]),
    ("nlfh_dmap_rho_vjp", ctypes.c_int, [_vp, _vp, ctypes.c_double, _vp, ctypes.c_int, _vp, _vp]),
    ("nlfh_fs_work_doubles", ctypes.c_longlong, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("nlfh_fs_prepare", ctypes.c_int, [_c_int_p, _vp, _c_double_p, ctypes.c_int, _vp, _vp]),
    ("nlfh_fs_project", ctypes.c_int, [_c_int_p, _vp, _vp, _vp]),
    ("nlfh_fs_macro_scatter", ctypes.c_int, [ctypes.c_int, _c_int_p, _vp, _vp, _vp]),
    ("nlfh_fs_macro_gather", ctypes.c_int, [ctypes.c_int, _c_int_p, _vp, _vp, _vp]),
    ("nlfh_fs_dot_part_doubles", ctypes.c_longlong, []),
    ("nlfh_fs_project_dot", ctypes.c_int, [_c_int_p, _vp, _vp, _vp, _vp]),
    ("nlfh_fs_cg_update_parts", ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_vp] * 11),
    ("nlfh_fs_cg_direction_parts", ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [_vp] * 10),
    ("nlfh_fs_cg_update", ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("nlfh_fs_cg_direction", ctypes.c_int, [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("nlfh_fs_toec", ctypes.c_int, [_c_int_p, _vp, _c_double_p, ctypes.c_int, _vp, _vp, _vp, _vp]),
    ("nlfh_fs_eval", ctypes.c_int, [_c_int_p, _vp, _c_double_p, ctypes.c_int, _vp, _c_double_p, _vp, _vp,
                                    ctypes.c_int, _vp, _vp, _vp, _vp]),
    ("nlfh_fs_noalloc_max_voxels", ctypes.c_longlong, []),
    ("nlfh_fs_eval_noalloc", ctypes.c_int, [_c_int_p, _vp, _c_double_p, ctypes.c_int, _vp, _c_double_p, _vp, _vp,
                                            ctypes.c_int, _vp, _vp, _vp]),
    ("nlfh_last_error", ctypes.c_char_p, []),
    ("nlfh_version", ctypes.c_int, []),
    ("nlfh_plan_create", ctypes.c_int, [_c_int_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                        ctypes.POINTER(_vp)]),
    ("nlfh_plan_destroy", ctypes.c_int, [_vp]),
    ("nlfh_plan_set_tiling", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("nlfh_plan_get_tiling", ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]),
    ("nlfh_plan_set_inner_tol", ctypes.c_int, [_vp, ctypes.c_double]),
    ("nlfh_plan_set_inner_restart", ctypes.c_int, [_vp, ctypes.c_int]),
    ("nlfh_plan_set_fft_tiling", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("nlfh_plan_get_fft_tiling", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp]),
    ("nlfh_stiffness_scale", ctypes.c_int, [_vp, _vp, ctypes.c_double, ctypes.c_double, _vp, _vp, _vp]),
    ("nlfh_apply_K", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    ("nlfh_load", ctypes.c_int, [_vp, _vp, _c_double_p, ctypes.c_int, _vp, _vp]),
    ("nlfh_residual", ctypes.c_int, [_vp, _vp, _c_double_p, _vp, ctypes.c_int, _vp, _vp]),
    ("nlfh_precond", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    ("nlfh_solve", ctypes.c_int, [_vp, _vp, _c_double_p, ctypes.c_int, _vp, _vp, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, _c_int_p, _c_int_p, _c_double_p, _vp]),
    ("nlfh_solve_rhs", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _c_int_p, _c_int_p, _c_double_p, _vp]),
    ("nlfh_ritz_start", ctypes.c_int, [_vp, _vp, _c_double_p, ctypes.c_int, _vp, _vp, _c_double_p, _vp]),
    ("nlfh_jacobi_bound", ctypes.c_double, [ctypes.c_double, ctypes.c_double]),
    ("nlfh_cheb_smooth", ctypes.c_int, [_vp, _vp, _c_double_p, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, _c_double_p, _vp]),
    ("nlfh_bilinear", ctypes.c_int, [_vp, _vp, _c_double_p, _vp, ctypes.c_int, _vp, _vp, _vp]),
    ("nlfh_project_mean", ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    ("nlfh_symbol", ctypes.c_int, [_vp, ctypes.c_int, _c_double_p]),
    ("nlfh_plan_bytes", ctypes.c_longlong, [_vp]),
    # learned warm start network kernels
    ("nlfh_net_conv", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                                     _vp]),
    ("nlfh_net_wgrad", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _vp, ctypes.c_longlong, _vp, _vp]),
    ("nlfh_net_act_grad", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int, _vp]),
    ("nlfh_net_bias_grad", ctypes.c_int, [_vp, _vp, ctypes.c_longlong, ctypes.c_int, ctypes.c_int, _vp,
                                          ctypes.c_longlong, _vp, _vp]),
    ("nlfh_net_pool", ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]),
    ("nlfh_net_upsum", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_float, ctypes.c_int, _vp]),
    ("nlfh_net_updist", ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                                       ctypes.c_int, _vp]),
    ("nlfh_net_slice_add", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_longlong, ctypes.c_int,
                                          ctypes.c_int, _vp]),
    ("nlfh_net_mse", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_longlong, _vp, ctypes.c_int, _vp]),
    ("nlfh_net_input", ctypes.c_int, [_vp, ctypes.c_longlong, _vp, _vp]),
    ("nlfh_net_readout", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_double, _vp,
                         ctypes.c_longlong, _vp, _vp]),
    ("nlfh_net_adam", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_longlong, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_float, _vp]),
    ("nlfh_net_f2bf", ctypes.c_int, [_vp, _vp, ctypes.c_longlong, _vp]),
    ("nlfh_net_wflip", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp]),
]

_lib = None


class NlfhError(RuntimeError):
    """A call into libnlfh.so returned a nonzero status."""


def load():
    """Load libnlfh.so (once). Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() or build_lib.py")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().nlfh_last_error().decode(errors="replace")
        raise NlfhError(f"libnlfh status {rc}: {msg}")
    return rc
