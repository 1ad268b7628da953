"""Host-only planner and input checks of libnnkifmm (CPU: no GPU call).

  * the nearest-map check run by nnk_tree_set_nearest before any of the map
    reaches the GPU: colptr[0] == 0 and monotone, rows in range, no null
    array that would be read;
  * the mixed-form M2L level plan: one merged job over all 316 offsets is
    chosen only where it walks fewer (tile, block) pairs, its table is
    built only for levels of at most 8 x 128 nodes, and a plan whose gather
    table holds an index outside [-1, nodeCount) is refused before any
    launch.
"""
import ctypes
import os

import numpy as np
import pytest

P = ctypes.POINTER


@pytest.fixture(scope="module")
def st(nnk):
    lib = ctypes.CDLL(os.path.join(os.path.dirname(nnk.library_path()), "libnnk_selftest.so"))
    lib.nnk_selftest_validate_nn.argtypes = [ctypes.c_int64, ctypes.c_int64, P(ctypes.c_double),
                                             P(ctypes.c_int64), P(ctypes.c_int64), P(ctypes.c_double),
                                             ctypes.c_char_p, ctypes.c_int]
    lib.nnk_selftest_m2l_plan.argtypes = [P(ctypes.c_int32), P(ctypes.c_int32), ctypes.c_int64, ctypes.c_int,
                                          P(ctypes.c_int32), P(ctypes.c_int64), ctypes.c_char_p, ctypes.c_int]
    return lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(P(t))


def _nn(st, ncol, nfine, colptr, row, val, fine=True):
    f = np.zeros((max(nfine, 1), 3)) if fine else None
    err = ctypes.create_string_buffer(256)
    rc = st.nnk_selftest_validate_nn(ncol, nfine, _p(f, ctypes.c_double), _p(colptr, ctypes.c_int64),
                                     _p(row, ctypes.c_int64), _p(val, ctypes.c_double), err, 256)
    return rc, err.value.decode()


def test_nearest_map_checks(st):
    cp = np.array([0, 2, 3, 5], np.int64)
    row = np.array([0, 4, 1, 2, 3], np.int64)
    val = np.array([1.0, 0.5, 1.0, 0.5, 1.0])
    assert _nn(st, 3, 5, cp, row, val) == (0, "")
    rc, msg = _nn(st, 3, 5, np.array([-1, 2, 3, 5], np.int64), row, val)
    assert rc == 1 and "colptr[0]" in msg
    rc, msg = _nn(st, 3, 5, np.array([0, 3, 2, 5], np.int64), row, val)
    assert rc == 1 and "monotone" in msg
    rc, msg = _nn(st, 3, 5, cp, None, val)
    assert rc == 1 and "null" in msg
    rc, msg = _nn(st, 3, 5, cp, row, None)
    assert rc == 1 and "null" in msg
    rc, msg = _nn(st, 3, 5, cp, np.array([0, 5, 1, 2, 3], np.int64), val)
    assert rc == 1 and "out of range" in msg
    rc, msg = _nn(st, 3, 5, cp, row, val, fine=False)
    assert rc == 1 and "fine" in msg
    assert _nn(st, 3, 0, np.zeros(4, np.int64), None, None, fine=False) == (0, "")   # empty map


def _plan(st, per_class, count, rng, fill=0.9, corrupt=None):
    sizes = np.full(8, per_class, np.int32)
    gv = np.where(rng.random(8 * 189 * per_class) < fill,
                  rng.integers(0, count, 8 * 189 * per_class), -1).astype(np.int32)
    if corrupt is not None:
        gv[len(gv) // 2] = corrupt
    merged = np.zeros(1, np.int32)
    pairs = np.zeros(2, np.int64)
    err = ctypes.create_string_buffer(256)
    rc = st.nnk_selftest_m2l_plan(_p(sizes, ctypes.c_int32), _p(gv, ctypes.c_int32), count, 128,
                                  _p(merged, ctypes.c_int32), _p(pairs, ctypes.c_int64), err, 256)
    return rc, int(merged[0]), pairs.tolist(), err.value.decode()


def test_m2l_plan_merges_small_levels_only(st):
    rng = np.random.default_rng(0)
    # 8 nodes per class (level 2 of a uniform tree): 8 x 189 class pairs
    # against 316 merged pairs -> merged
    rc, merged, (cp, mp), _ = _plan(st, 8, 73, rng)
    assert rc == 0 and merged == 1 and cp == 8 * 189 and mp == 316
    # 128 per class (1,024 nodes, the limit): the table is built, but each
    # 128-column tile of the merged job holds one class (its 189 offsets):
    # a tie, and the class jobs are kept
    rc, merged, (cp, mp), _ = _plan(st, 128, 5000, rng)
    assert rc == 0 and merged == 0 and cp == 8 * 189 and mp == 8 * 189
    # 129 per class: above the limit, the merged table is not even built
    rc, merged, (cp, mp), _ = _plan(st, 129, 5000, rng)
    assert rc == 0 and merged == 0 and mp == -1 and 8 * 189 < cp <= 8 * 2 * 189


@pytest.mark.parametrize("bad", [-2, 73, 1 << 30])
def test_m2l_plan_refuses_out_of_range_gathers(st, bad):
    rng = np.random.default_rng(1)
    rc, _, _, msg = _plan(st, 8, 73, rng, corrupt=bad)
    assert rc == 1 and "gather index out of range" in msg, msg
    rc, _, _, msg = _plan(st, 200, 5000, rng, corrupt=5000 if bad == 73 else bad)
    assert rc == 1 and "gather index out of range" in msg, msg
