"""Oracle pins for the octree and the U/V/W/X lists (CPU).

Independent checks: a hand-worked split, the one-sphere tree (root + 8
leaves), structural invariants, list sizes of the complete uniform tree
computed from grid geometry alone, list symmetries, and interaction
coverage: every (target leaf, source leaf) pair is accounted for exactly
once by U, W, and the V/X lists of the target's ancestors -- the property
that makes the FMM equal to the direct sum up to the expansion error.
"""
import collections
import itertools

import numpy as np
import pytest

from geometry import blobs, cube_sphere, grid16


def test_hand_worked_split(oracle):
    # root cube [0,1]; centre 0.5; points on the centre go low (strict >)
    pts = np.array([[0, 0, 0], [1, 1, 1], [0.5, 0.5, 0.5], [0.6, 0.4, 0.6], [0.2, 0.9, 0.1],
                    [0.9, 0.1, 0.2], [0.4, 0.4, 0.9], [0.7, 0.8, 0.3], [0.1, 0.2, 0.3]], float)
    t = oracle.Tree(pts, pts, nodeCapacity=8)   # capacity = min(8, 9-1) = 8: root splits once
    assert t.nodeCount == 9
    np.testing.assert_array_equal(t.nodeCorners[0], [0, 0, 0, 1, 1, 1])
    # child k (1-based) = 1 + 4[x>c] + 2[y>c] + [z>c]; ids 2..9
    expect = [1 + 1 + 4 * (p[0] > .5) + 2 * (p[1] > .5) + (p[2] > .5) for p in pts]
    np.testing.assert_array_equal(t.pointIndex, expect)
    np.testing.assert_array_equal(t.nodeCorners[1 + 5 - 1], [0.5, 0, 0, 1, 0.5, 0.5])  # child 5: x high
    np.testing.assert_array_equal(t.nodeCorners[1 + 2 - 1], [0, 0, 0.5, 0.5, 0.5, 1])  # child 2: z high
    np.testing.assert_array_equal(t.nodeLevel, [0] + [1] * 8)
    np.testing.assert_array_equal(t.nodeParents, [0] + [1] * 8)
    np.testing.assert_array_equal(t.nodeChildren[0], np.arange(2, 10))


def test_dfs_creation_numbering(oracle):
    # refine only the low octant: its children get ids 10..17 before any
    # other split can happen
    rng = np.random.default_rng(4)
    p = np.concatenate([rng.random((20, 3)) * 0.5, rng.random((5, 3)) * 0.5 + 0.5, [[0, 0, 0], [1, 1, 1]]])
    t = oracle.Tree(p, p, nodeCapacity=10)
    np.testing.assert_array_equal(t.nodeChildren[1], np.arange(10, 18))
    assert np.all(t.nodeLevel[9:17] == 2)


def test_sphere_tree_is_root_plus_eight(oracle):
    p = cube_sphere(8)
    t = oracle.Tree(p, p, nodeCapacity=64)
    assert t.nodeCount == 9
    assert np.all(np.bincount(t.pointIndex)[2:] == 48)
    for i in range(2, 10):   # all leaves mutually adjacent
        np.testing.assert_array_equal(t.interactions(i, 0), np.arange(2, 10))
        for k in (1, 2, 3):
            assert len(t.interactions(i, k)) == 0


def test_max_depth_error(oracle):
    p = np.concatenate([np.zeros((1, 3)), np.ones((1, 3)), 1e-9 * np.random.default_rng(0).random((30, 3))])
    with pytest.raises(RuntimeError, match="Max level reached"):
        oracle.Tree(p, p, nodeCapacity=4, maxDepth=6)


@pytest.mark.parametrize("case", ["uniform", "blobs", "distinct"])
def test_structural_invariants(oracle, case):
    rng = np.random.default_rng(5)
    if case == "uniform":
        p = t = rng.random((3000, 3)); cap = 32
    elif case == "blobs":
        p = t = blobs(rng); cap = 16
    else:
        p = rng.random((2000, 3)); t = np.concatenate([rng.random((1500, 3)), p[:300]]); cap = 25
    tr = oracle.Tree(p, t, nodeCapacity=cap)
    C = tr.nodeCorners
    # each point lies in its leaf's closed box; leaves hold <= capacity
    lo, hi = C[tr.pointIndex - 1, :3], C[tr.pointIndex - 1, 3:]
    assert np.all(tr.points >= lo) and np.all(tr.points <= hi)
    assert np.all(tr.nodeChildren[tr.pointIndex - 1, 0] == 0)
    assert np.bincount(tr.pointIndex).max() <= tr.nodeCapacity
    for n in range(1, tr.nodeCount + 1):
        ch = tr.nodeChildren[n - 1]
        if ch[0] == 0:
            continue
        np.testing.assert_array_equal(ch, np.arange(ch[0], ch[0] + 8))   # consecutive ids
        assert np.all(tr.nodeParents[ch - 1] == n)
        c = C[n - 1]
        mid = (c[:3] + c[3:]) / 2
        for k, cid in enumerate(ch):
            g = np.array([(k >> 2) & 1, (k >> 1) & 1, k & 1])
            np.testing.assert_array_equal(C[cid - 1, :3], np.where(g, mid, c[:3]))
            np.testing.assert_array_equal(C[cid - 1, 3:], np.where(g, c[3:], mid))
    # points: unique rows of [pot; tgt] in first-appearance order
    allp = np.concatenate([p, t])
    _, first = np.unique(allp, axis=0, return_index=True)
    np.testing.assert_array_equal(tr.points, allp[np.sort(first)])


def _grid_lists(level):
    """|U| per leaf cell and |V| per cell of the complete tree, from integer
    cell coordinates only."""
    n = 2 ** level
    cells = np.array(list(itertools.product(range(n), repeat=3)))
    par = cells // 2
    U = np.zeros(len(cells), int)
    V = np.zeros(len(cells), int)
    for i, c in enumerate(cells):
        d = np.abs(cells - c).max(1)
        dp = np.abs(par - par[i]).max(1)
        U[i] = (d <= 1).sum()
        V[i] = ((dp <= 1) & (d >= 2)).sum()
    return cells, U, V


def test_complete_uniform_tree_list_sizes(oracle):
    p = grid16()
    t = oracle.Tree(p, p, nodeCapacity=1)
    assert t.nodeCount == 1 + 8 + 64 + 512 + 4096
    vtot = collections.Counter()
    for n in range(2, t.nodeCount + 1):
        vtot[t.nodeLevel[n - 1]] += len(t.interactions(n, 1))
    # per-level V-pair totals of a complete octree (levels 2-4)
    assert vtot[2] == 3096 and vtot[3] == 53352 and vtot[4] == 584136
    for lev in (2, 3, 4):
        cells, U, V = _grid_lists(lev)
        assert V.sum() == vtot[lev]
    cells, U, V = _grid_lists(4)
    assert U.max() == 27 and V.max() == 189
    leaves = np.where(t.nodeLevel == 4)[0] + 1
    us = sorted(len(t.interactions(n, 0)) for n in leaves)
    assert us == sorted(U.tolist())
    assert all(len(t.interactions(n, 2)) == 0 and len(t.interactions(n, 3)) == 0 for n in leaves)


@pytest.mark.parametrize("case", ["blobs", "distinct", "shell"])
def test_list_symmetry_and_coverage(oracle, case):
    rng = np.random.default_rng(6)
    if case == "blobs":
        p = t = blobs(rng); cap = 16
    elif case == "distinct":
        p = rng.random((2500, 3)); t = np.concatenate([rng.random((1500, 3)), p[:500]]); cap = 20
    else:
        p = t = cube_sphere(24) * 0.9 + 0.05 * rng.random((6 * 24 * 24, 3)); cap = 24
    tr = oracle.Tree(p, t, nodeCapacity=cap)
    nc = tr.nodeCount
    ch = tr.nodeChildren
    pl = set(tr.pointIndex[tr.potentials].tolist())
    tl = set(tr.pointIndex[tr.targets].tolist())
    U = {n: set(tr.interactions(n, 0).tolist()) for n in range(1, nc + 1)}
    V = {n: set(tr.interactions(n, 1).tolist()) for n in range(1, nc + 1)}
    W = {n: tr.interactions(n, 2).tolist() for n in range(1, nc + 1)}
    X = {n: tr.interactions(n, 3).tolist() for n in range(1, nc + 1)}
    for a in range(2, nc + 1):
        for b in V[a]:
            assert a in V[b]                          # V symmetric
        for b in U[a]:
            if b in tl and a in pl:
                assert a in U[b]                      # U symmetric among source+target leaves
        for w in W[a]:
            assert a in X[w]                          # X = W^T
        for x in X[a]:
            assert a in W[x]
    def leaves_under(n):
        st, out = [n], []
        while st:
            y = st.pop()
            if ch[y - 1, 0] == 0:
                out.append(y)
            else:
                st.extend(ch[y - 1].tolist())
        return out
    under = {n: [l for l in leaves_under(n) if l in pl] for n in range(1, nc + 1)}
    for i in tl:
        cnt = collections.Counter()
        for j in U[i]:
            cnt[j] += 1
        for w in W[i]:
            for j in under[w]:
                cnt[j] += 1
        a = i
        while a > 1:
            for v in V[a]:
                for j in under[v]:
                    cnt[j] += 1
            for x in X[a]:
                if x in pl:
                    cnt[x] += 1
            a = tr.nodeParents[a - 1]
        assert set(cnt) == pl and set(cnt.values()) == {1}, (i, collections.Counter(cnt.values()))


@pytest.mark.parametrize("case", ["A", "B"])
def test_hand_worked_w_and_x_lists(oracle, case):
    """W coarsening order (re-sorted rest + last merged node, levels bottom
    up) and X's column-major find, against lists worked out by hand
    (tests/hand_trees.py gives the derivations)."""
    import hand_trees
    make, expect = hand_trees.CASES[case]
    p = make()
    t = oracle.Tree(p, p, nodeCapacity=1)
    hand_trees.check(t.interactions, t.nodeCount, expect)
