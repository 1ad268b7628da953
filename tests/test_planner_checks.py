"""Host-only planner and input checks of libnnkifmm (CPU: no GPU call).

  * the nearest-map check run by nnk_tree_set_nearest before any of the map
    reaches the GPU: colptr[0] == 0 and monotone, rows in range, no null
    array that would be read;
  * the mixed-form M2L level plan: one merged job over all 316 offsets is
    chosen only where it walks fewer (tile, block) pairs, its table is
    built only for levels of at most 8 x 128 nodes, and a plan whose gather
    table holds an index outside [-1, nodeCount) is refused before any
    launch;
  * the FP64 forms' per-offset M2L plan (sparse adaptive levels): chosen
    only where its padded work is below the class tiles', every V-list
    pair exactly once, jobs of at most 64 pairs of one offset laid out as
    contiguous runs per XCD (job 8k + x), per-target lists ordered by
    offset slot, out-of-range gathers refused.
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
    i32, i64 = P(ctypes.c_int32), ctypes.c_int64
    lib.nnk_selftest_m2l_pairs.argtypes = [i32, i32, i64, ctypes.c_int, i64, i64, P(ctypes.c_int64), i32, i32, i32,
                                           i64, i32, i32, i64, i32, i32, i64, ctypes.c_char_p, ctypes.c_int]
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


def _slot(ox, oy, oz):
    return (ox + 3) * 49 + (oy + 3) * 7 + (oz + 3)


def _visit_paths():
    """Per parity class p = 4 (ox mod 2) + 2 (oy mod 2) + (oz mod 2), the
    boustrophedon path over its offsets in [-3, 3]^3: x ascending, y
    ascending on even x steps and descending on odd ones, z ascending on
    even rows and descending on odd ones (rows counted through the class)."""
    paths = []
    for p in range(8):
        X, Y, Z = ([v for v in range(-3, 4) if v & 1 == (p >> s) & 1] for s in (2, 1, 0))
        path, row = [], 0
        for i, x in enumerate(X):
            for y in (Y if i % 2 == 0 else Y[::-1]):
                path += [(x, y, z) for z in (Z if row % 2 == 0 else Z[::-1])]
                row += 1
        paths.append(path)
    return paths


def _visit_order():
    """The M2L visiting order of the 316 V-list offsets (as slots): the
    parity classes in turn, each along its path, near offsets (max |o| < 2)
    left out."""
    return [_slot(*o) for path in _visit_paths() for o in path if max(map(abs, o)) >= 2]


def _class_slots(a):
    """The V-list offsets of a child of class a (bits x, y, z) as slots of
    the 7^3 offset table, in block order: the visiting order restricted to
    the offsets within [-2 - g, 3 - g] per axis (g the class bit)."""
    g = ((a >> 2) & 1, (a >> 1) & 1, a & 1)
    out = []
    for sl in _visit_order():
        o = (sl // 49 - 3, (sl // 7) % 7 - 3, sl % 7 - 3)
        if all(-2 - g[d] <= o[d] <= 3 - g[d] for d in range(3)):
            out.append(sl)
    return out


def test_visit_order_is_a_path_per_parity_class():
    order = _visit_order()
    # every non-adjacent offset of the 7^3 table once
    assert sorted(order) == sorted(_slot(x, y, z) for x in range(-3, 4) for y in range(-3, 4)
                                   for z in range(-3, 4) if max(abs(x), abs(y), abs(z)) >= 2)
    # within a class, consecutive offsets of the path differ by 2 in one coordinate
    for path in _visit_paths():
        for u, v in zip(path, path[1:]):
            d = sorted(abs(u[k] - v[k]) for k in range(3))
            assert d == [0, 0, 2], (u, v)
    # each child class sees 189 offsets
    assert all(len(_class_slots(a)) == 189 for a in range(8))


def _pairs(st, sizes, gv, count, M=456, cap=1 << 28):
    sizes = np.asarray(sizes, np.int32)
    info = np.zeros(5, np.int64)
    nj, npair, nn = 8 * 4096, max(1, int((gv >= 0).sum())), int(sizes.sum())
    js, jp, jn = (np.zeros(nj, np.int32) for _ in range(3))
    src, idx = np.zeros(npair, np.int32), np.zeros(npair, np.int32)
    node, ptr = np.zeros(nn, np.int32), np.zeros(nn + 1, np.int32)
    err = ctypes.create_string_buffer(256)
    i32 = ctypes.c_int32
    rc = st.nnk_selftest_m2l_pairs(_p(sizes, i32), _p(gv, i32), count, 64, M, cap, _p(info, ctypes.c_int64),
                                   _p(js, i32), _p(jp, i32), _p(jn, i32), nj, _p(src, i32), _p(idx, i32), npair,
                                   _p(node, i32), _p(ptr, i32), nn, err, 256)
    return rc, info, (js, jp, jn, src, idx, node, ptr), err.value.decode()


def _sparse_level(rng, sizes, fill, count):
    return np.concatenate([np.where(rng.random(189 * n) < fill, rng.integers(0, count, 189 * n), -1)
                           for n in sizes]).astype(np.int32)


def test_m2l_pair_plan_tables(st):
    rng = np.random.default_rng(3)
    sizes = [40, 7, 90, 0, 33, 64, 5, 120]
    count = 2000
    gv = _sparse_level(rng, sizes, 0.08, count)
    rc, info, (js, jp, jn, src, idx, node, ptr), msg = _pairs(st, sizes, gv, count)
    assert rc == 0, msg
    use, cw, pw, npairs, nj = info.tolist()
    assert use == 1 and pw < cw and npairs == int((gv >= 0).sum()) and nj % 8 == 0
    # expected pairs: (slot, target, source) from the gather tables
    want, off, base = [], 0, 0
    for a, n in enumerate(sizes):
        sl = _class_slots(a)
        g = gv[off:off + 189 * n].reshape(189, n)
        for b in range(189):
            for c in range(n):
                if g[b, c] >= 0:
                    want.append((sl[b], base + c, int(g[b, c])))
        off += 189 * n
        base += n
    # jobs: at most 64 pairs of one slot; tiles in visiting order, contiguous
    # runs per XCD (job 8k + x), and together they cover pairs 0..npairs-1
    js, jp, jn = js[:nj], jp[:nj], jn[:nj]
    assert (jn <= 64).all() and ((jn == 0) == (js == -1)).all()
    seq = [(int(js[8 * k + x]), int(jp[8 * k + x]), int(jn[8 * k + x]))
           for x in range(8) for k in range(nj // 8) if jn[8 * k + x] > 0]
    assert [t[1] for t in seq] == sorted(t[1] for t in seq)
    covered = np.concatenate([np.arange(p0, p0 + n) for _, p0, n in seq])
    assert np.array_equal(covered, np.arange(npairs))
    rank = {sl: i for i, sl in enumerate(_visit_order())}
    assert [rank[t[0]] for t in seq] == sorted(rank[t[0]] for t in seq)
    runs = [sum(1 for k in range(nj // 8) if jn[8 * k + x] > 0) for x in range(8)]
    assert max(runs) - min(runs) <= 1
    pair_slot = np.empty(npairs, np.int64)
    for sl, p0, n in seq:
        pair_slot[p0:p0 + n] = sl
    # per target: its pairs, in the visiting order; together every pair once
    assert sorted(node.tolist()) == list(range(sum(sizes)))
    got = []
    for j, t in enumerate(node.tolist()):
        ps = idx[ptr[j]:ptr[j + 1]]
        assert [rank[int(pair_slot[p])] for p in ps] == sorted(rank[int(pair_slot[p])] for p in ps)
        got += [(int(pair_slot[p]), t, int(src[p])) for p in ps]
    assert sorted(got) == sorted(want)
    assert np.array_equal(np.sort(idx), np.arange(npairs))


def test_m2l_pair_plan_choice(st):
    rng = np.random.default_rng(4)
    # a full uniform level: the class tiles walk no empty column
    sizes = [64] * 8
    gv = rng.integers(0, 600, 189 * 64 * 8).astype(np.int32)
    rc, info, _, msg = _pairs(st, sizes, gv, 600)
    assert rc == 0 and info[0] == 0 and info[1] == 8 * 189 * 64, msg
    # sparse, but the pair buffer would pass its limit: class tiles
    gv = _sparse_level(rng, sizes, 0.05, 600)
    rc, info, _, msg = _pairs(st, sizes, gv, 600, M=456, cap=456 * 100)
    assert rc == 0 and info[0] == 0 and info[3] > 100, msg
    rc, info, _, msg = _pairs(st, sizes, gv, 600)
    assert rc == 0 and info[0] == 1, msg


@pytest.mark.parametrize("bad", [-2, 600])
def test_m2l_pair_plan_refuses_out_of_range_gathers(st, bad):
    rng = np.random.default_rng(5)
    gv = _sparse_level(rng, [30] * 8, 0.1, 600)
    gv[len(gv) // 3] = bad
    rc, _, _, msg = _pairs(st, [30] * 8, gv, 600)
    assert rc == 1 and "gather index out of range" in msg, msg
