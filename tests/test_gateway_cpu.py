"""The MATLAB marshalling layer (stokes-nnkifmm_amd/matlab/nnk_gateway.c)
on the CPU: every function its header declares is exported, the layout
helpers turn row-major into MATLAB column-major and mwIndex into the ABI's
integer types with range checks, and unknown or released handle ids are
refused with a message instead of being dereferenced.  (The calls that
reach the GPU are in tests/test_gpu_gateway.py.)"""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
P = ctypes.POINTER


@pytest.fixture(scope="module")
def gw(nnk):
    lib = ctypes.CDLL(os.path.join(os.path.dirname(nnk.library_path()), "libnnk_gateway.so"))
    lib.nnkgw_last_error.restype = ctypes.c_char_p
    lib.nnkgw_live.restype = ctypes.c_int64
    return lib


def test_gateway_exports_its_header(gw):
    with open(os.path.join(ROOT, "stokes-nnkifmm_amd", "matlab", "nnk_gateway.h")) as fh:
        names = set(re.findall(r"\b(nnkgw_[a-z0-9_]+)\s*\(", fh.read()))
    assert len(names) >= 20
    assert [n for n in sorted(names) if not hasattr(gw, n)] == []


def test_layout_helpers(gw):
    a = np.arange(12, dtype=np.float64).reshape(3, 4)          # row-major 3 x 4
    out = np.zeros(12)
    gw.nnkgw_rowmajor_to_colmajor(a.ctypes.data_as(P(ctypes.c_double)), ctypes.c_int64(3), ctypes.c_int64(4),
                                  out.ctypes.data_as(P(ctypes.c_double)))
    np.testing.assert_array_equal(out, a.ravel(order="F"))
    idx = np.array([0, 5, 2 ** 31 - 1], np.uint64)
    o32 = np.zeros(3, np.int32)
    assert gw.nnkgw_index_to_i32(idx.ctypes.data_as(P(ctypes.c_size_t)), ctypes.c_int64(3),
                                 o32.ctypes.data_as(P(ctypes.c_int32))) == 0
    np.testing.assert_array_equal(o32, [0, 5, 2 ** 31 - 1])
    idx[1] = 2 ** 31
    assert gw.nnkgw_index_to_i32(idx.ctypes.data_as(P(ctypes.c_size_t)), ctypes.c_int64(3),
                                 o32.ctypes.data_as(P(ctypes.c_int32))) == 1
    assert b"out of range" in gw.nnkgw_last_error()
    o64 = np.zeros(3, np.int64)
    assert gw.nnkgw_index_to_i64(idx.ctypes.data_as(P(ctypes.c_size_t)), ctypes.c_int64(3),
                                 o64.ctypes.data_as(P(ctypes.c_int64))) == 0
    np.testing.assert_array_equal(o64, idx.astype(np.int64))


def test_unknown_handles_are_refused(gw):
    assert gw.nnkgw_live(0) == 0
    assert gw.nnkgw_free(ctypes.c_uint64(7)) == 1
    assert b"unknown or released" in gw.nnkgw_last_error()
    u = np.zeros(3)
    assert gw.nnkgw_fmm_apply(ctypes.c_uint64(0), 0, u.ctypes.data_as(P(ctypes.c_double)),
                              u.ctypes.data_as(P(ctypes.c_double))) == 1
    assert b"not a live KIFMM handle" in gw.nnkgw_last_error()
    n = ctypes.c_int64()
    assert gw.nnkgw_mobility_size(ctypes.c_uint64(3), ctypes.byref(n)) == 1
    assert gw.nnkgw_tree_sizes(ctypes.c_uint64(1), None, None, None, None, None) == 1
    gw.nnkgw_free_all()
    assert gw.nnkgw_live(0) == 0
