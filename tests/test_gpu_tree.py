"""The octree (subdivision on the GPU) and its interaction lists against
the oracle (GPU).

The tree and its interaction lists must be bit-exact: same ids, corners,
point-to-leaf map and list contents in the same order.  Also: error paths.
"""
import os

import numpy as np
import pytest

from geometry import blobs, cube_sphere, grid16

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
pytestmark = pytest.mark.gpu


def _cases():
    rng = np.random.default_rng(11)
    p = rng.random((4000, 3))
    yield "uniform", p, p, 32
    b = blobs(rng)
    yield "blobs", b, b, 12
    s = cube_sphere(24) + 0.01 * rng.standard_normal((6 * 24 * 24, 3))
    yield "shell", s, s, 20
    q = rng.random((2500, 3))
    yield "distinct", q, np.concatenate([rng.random((1500, 3)), q[:400]]), 30
    d = np.concatenate([p[:500], p[:200]])   # duplicated rows
    yield "duplicates", d, d, 16
    g = grid16()
    yield "complete", g, g, 1
    c = cube_sphere(8)
    yield "sphere", c, c, 64


@pytest.mark.parametrize("name,pot,tgt,cap", list(_cases()), ids=[c[0] for c in _cases()])
def test_tree_bit_exact(oracle, nnk, name, pot, tgt, cap):
    a = nnk.OcTree(pot, tgt, "nodeCapacity", cap)
    b = oracle.Tree(pot, tgt, nodeCapacity=cap)
    assert a.nodeCount == b.nodeCount and a.nodeCapacity == b.nodeCapacity
    for attr in ("points", "potentials", "targets", "pointIndex", "nodeCorners", "nodeLevel",
                 "nodeParents", "nodeChildren"):
        np.testing.assert_array_equal(getattr(a, attr), getattr(b, attr), err_msg=attr)
    for kind in range(4):
        for n in range(1, a.nodeCount + 1):
            np.testing.assert_array_equal(a.interactions(n, kind + 1), b.interactions(n, kind),
                                          err_msg=f"list {'UVWX'[kind]} of node {n}")


def test_max_depth_and_bad_input(nnk):
    p = np.concatenate([np.zeros((1, 3)), np.ones((1, 3)), 1e-9 * np.random.default_rng(0).random((30, 3))])
    with pytest.raises(nnk.NNKError, match="Max level reached"):
        nnk.OcTree(p, p, "nodeCapacity", 4, "maxDepth", 6)
    bad = np.random.default_rng(1).random((10, 3))
    bad[3, 1] = np.nan
    with pytest.raises(nnk.NNKError, match="NaN"):
        nnk.OcTree(bad, bad)
    with pytest.raises(nnk.NNKError):
        nnk.OcTree(bad[:1] * 0, bad[:1] * 0)   # a single distinct point
    with pytest.raises(ValueError):
        nnk.OcTree(p, p, "nodeCapacity")        # unpaired option
    with pytest.raises(ValueError):
        nnk.OcTree(p, p, "noSuchOption", 3)


def test_options_name_value_and_keywords(nnk):
    p = np.random.default_rng(2).random((500, 3))
    a = nnk.OcTree(p, p, "nodeCapacity", 20, "maxDepth", 10)
    b = nnk.OcTree(p, p, nodeCapacity=20, maxDepth=10)
    assert a.nodeCount == b.nodeCount and a.arguments["nodeCapacity"] == 20
    np.testing.assert_array_equal(a.getPotentialPoints(), p)
    np.testing.assert_array_equal(a.getTargetPoints(), p)


def test_root_cube_option(oracle, nnk):
    """'rootCube' (an extension: fixed root box for time stepping).  With the
    points' own bounding cube it reproduces the default tree; with a larger
    cube the product tree equals the oracle's and every node has width
    (hi - lo) / 2^level and contains its points."""
    rng = np.random.default_rng(5)
    p = rng.random((3000, 3)) * [0.7, 0.9, 0.5] + 0.1
    m, M = p.min(), p.max()
    a = nnk.OcTree(p, p, "nodeCapacity", 24)
    b = nnk.OcTree(p, p, "nodeCapacity", 24, "rootCube", [m, M])
    for attr in ("points", "pointIndex", "nodeCorners", "nodeLevel", "nodeParents", "nodeChildren"):
        np.testing.assert_array_equal(getattr(a, attr), getattr(b, attr), err_msg=attr)
    lo, hi = -1.0, 2.0
    c = nnk.OcTree(p, p, "nodeCapacity", 24, "rootCube", [lo, hi])
    d = oracle.Tree(p, p, nodeCapacity=24, rootCube=[lo, hi])
    assert c.nodeCount == d.nodeCount
    for attr in ("points", "pointIndex", "nodeCorners", "nodeLevel", "nodeParents", "nodeChildren"):
        np.testing.assert_array_equal(getattr(c, attr), getattr(d, attr), err_msg=attr)
    for kind in range(4):
        for n in range(1, c.nodeCount + 1):
            np.testing.assert_array_equal(c.interactions(n, kind + 1), d.interactions(n, kind))
    w = c.nodeCorners[:, 3:] - c.nodeCorners[:, :3]
    np.testing.assert_allclose(w, ((hi - lo) / 2.0 ** c.nodeLevel)[:, None] * np.ones(3), rtol=1e-14)
    box = c.nodeCorners[c.pointIndex - 1]
    assert np.all(c.points >= box[:, :3]) and np.all(c.points <= box[:, 3:])
    with pytest.raises(nnk.NNKError, match="root cube"):
        nnk.OcTree(p, p, "rootCube", [0.2, 2.0])
    with pytest.raises(RuntimeError, match="root cube"):
        oracle.Tree(p, p, rootCube=[0.2, 2.0])


@pytest.mark.parametrize("case", ["A", "B"])
def test_hand_worked_w_and_x_lists(nnk, case):
    """The GPU list build against W/X lists worked out by hand from
    OcTree.m:279-288,343-360 (derivations in tests/hand_trees.py)."""
    import hand_trees
    make, expect = hand_trees.CASES[case]
    p = make()
    t = nnk.OcTree(p, p, "nodeCapacity", 1)
    hand_trees.check(lambda n, k: t.interactions(n, k + 1), t.nodeCount, expect)
