"""ctypes binding of libnnkifmm.so (declarations in include/nnkifmm.h).

The library is the only compute path: if it is missing or fails to load,
every call raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libnnkifmm.so")

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_vp = ctypes.c_void_p
_h = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); mirrors include/nnkifmm.h
SIGNATURES = {
    "nnk_version": (ctypes.c_int32, []),
    "nnk_last_error": (ctypes.c_char_p, []),
    "nnk_tree_create": (ctypes.c_int32, [_d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int32,
                                         ctypes.c_int32, _h]),
    "nnk_tree_create_in_cube": (ctypes.c_int32, [_d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int32,
                                                 ctypes.c_int32, _d, _h]),
    "nnk_tree_set_nearest": (ctypes.c_int32, [_vp, _d, ctypes.c_int64, _i64, _i64, _d]),
    "nnk_tree_build_times": (ctypes.c_int32, [_vp, _d]),
    "nnk_tree_sizes": (ctypes.c_int32, [_vp, _i64, _i64, _i64, _i32, _i32]),
    "nnk_tree_export": (ctypes.c_int32, [_vp, _d, _u8, _u8, _i32, _d, _i32, _i32, _i32]),
    "nnk_tree_list_size": (ctypes.c_int32, [_vp, ctypes.c_int32, _i64]),
    "nnk_tree_list_export": (ctypes.c_int32, [_vp, ctypes.c_int32, _i64, _i32]),
    "nnk_tree_destroy": (None, [_vp]),
    "nnk_fmm_create": (ctypes.c_int32, [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _h]),
    "nnk_fmm_apply": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_fmm_apply_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_fmm_reduced_apply": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_fmm_reduced_apply_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_fmm_destroy": (None, [_vp]),
    "nnk_direct_apply": (ctypes.c_int32, [_d, ctypes.c_int64, _d, _d, ctypes.c_int64, ctypes.c_double,
                                          ctypes.c_double, _d]),
    "nnk_direct_apply_device": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, _vp, ctypes.c_int64,
                                                 ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "nnk_direct_apply_device_noalloc": (ctypes.c_int32, [_vp, ctypes.c_int64, _vp, _vp, ctypes.c_int64,
                                                         ctypes.c_double, ctypes.c_double, _vp, _vp]),
    "nnk_fmm_set_profiling": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_fmm_phase_times": (ctypes.c_int32, [_vp, _d, _i64, ctypes.c_int32]),
    "nnk_fmm_stats": (ctypes.c_int32, [_vp, _i64, ctypes.c_int32]),
    "nnk_fmm_reduced_solve": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_fmm_reduced_solve_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_nearest_matrix": (ctypes.c_int32, [_d, ctypes.c_int64, _d, ctypes.c_int64, _i64, _i64, _d,
                                            ctypes.c_int64, _i64]),
    "nnk_mobility_create": (ctypes.c_int32, [_vp, _vp, ctypes.c_int64, _i64, _i32, _d, _i64, _i32, _d,
                                             _d, _d, _h]),
    "nnk_mobility_apply": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_mobility_apply_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_mobility_precon": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_mobility_precon_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_mobility_solve": (ctypes.c_int32, [_vp, _d, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _d,
                                            _i32, _d, _i32, _d]),
    "nnk_mobility_solve_refined": (ctypes.c_int32, [_vp, _vp, _vp, _d, ctypes.c_double, ctypes.c_int32,
                                                    ctypes.c_double, ctypes.c_int32, _d, _i32, _d, _i32,
                                                    _i32]),
    "nnk_mobility_solve_fgmres": (ctypes.c_int32, [_vp, _vp, _vp, _d, ctypes.c_double, ctypes.c_int32,
                                                    ctypes.c_double, ctypes.c_int32, _d, _i32, _d, _i32,
                                                    _i32]),
    "nnk_fmm_set_graph": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_fmm_set_autotune": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_fmm_launch_config": (ctypes.c_int32, [_vp, _i32, _i32, ctypes.c_int32]),
    "nnk_fmm_set_launch_config": (ctypes.c_int32, [_vp, _i32, _i32, ctypes.c_int32]),
    "nnk_fmm_leaf_threads": (ctypes.c_int32, [_vp, _i32]),
    "nnk_op_cache_limit": (ctypes.c_int32, [ctypes.c_int64]),
    "nnk_op_cache_stats": (ctypes.c_int32, [_i64, _i64, _i64, _i64]),
    "nnk_mobility_set_graph": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_mobility_destroy": (None, [_vp]),
    "nnk_batch_create": (ctypes.c_int32, [ctypes.c_int32, _i64, _d, _i64, _d, _i64, _i32, _d, _d, _d, _d, _i32,
                                          _i64, _d, ctypes.c_double, ctypes.c_int32, _vp]),
    "nnk_batch_size": (ctypes.c_int32, [_vp, _i64]),
    "nnk_batch_apply": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_batch_apply_transpose": (ctypes.c_int32, [_vp, _d, _d]),
    "nnk_batch_apply_device": (ctypes.c_int32, [_vp, _vp, _vp, _vp]),
    "nnk_batch_set_graph": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_batch_set_autotune": (ctypes.c_int32, [_vp, ctypes.c_int32]),
    "nnk_batch_p2p_waves": (ctypes.c_int32, [_vp, _i32]),
    "nnk_batch_solve": (ctypes.c_int32, [_vp, _d, ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _d, _i32, _d,
                                         _i32]),
    "nnk_batch_gradient": (ctypes.c_int32, [_vp, _d, _d, _d, _d, _d]),
    "nnk_batch_destroy": (None, [_vp]),
}

_lib = None


class NNKError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[nnkifmm status {code}] {msg}")
        self.code = code


def lib():
    """Load libnnkifmm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime under the same SONAME as ROCm's;
        # load it first so the process holds a single runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                               "or python stokes-nnkifmm_amd/build.py")
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise NNKError(rc, lib().nnk_last_error().decode())


def ptr(a, t=_d):
    return a.ctypes.data_as(t)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
