"""Oracle pins for the KIFMM passes, the NN map and the reduced operators.

  * root + 8 mutually adjacent leaves: U = all leaves and V = W = X = empty,
    so the FMM equals the direct sum up to summation order;
  * deeper trees: error against the direct sum is small and falls with the
    corona resolution (4 -> 6 -> 8);
  * nearestMatrix against a k-d tree, equal weights on exact ties;
  * reduced pass = the block-diagonal direct sum; reduced solve inverts it.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.spatial import cKDTree

from geometry import blobs, cube_sphere


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_root_plus_eight_equals_direct(oracle):
    rng = np.random.default_rng(0)
    p = cube_sphere(8)
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=64)
    u = t.matvec(q, 0.01, 1.0, 6, 1)
    assert rel(u, oracle.direct(p, q, p, 0.01)) < 1e-14


def test_error_falls_with_corona_resolution(oracle):
    rng = np.random.default_rng(1)
    p = rng.random((2500, 3))
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=40)
    assert t.nodeLevel.max() >= 2
    ud = oracle.direct(p, q, p, 1e-3)
    e4, e6 = rel(t.matvec(q, 1e-3, 1.0, 4, 1), ud), rel(t.matvec(q, 1e-3, 1.0, 6, 1), ud)
    assert e6 < e4 / 5 and e6 < 1e-4, (e4, e6)


def test_adaptive_tree_accuracy(oracle):
    rng = np.random.default_rng(2)
    p = blobs(rng)
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=32)
    assert sum(len(t.interactions(n, 2)) for n in range(2, t.nodeCount + 1)) > 0   # W/X exercised
    assert rel(t.matvec(q, 5e-3, 1.0, 4, 1), oracle.direct(p, q, p, 5e-3)) < 5e-3


def test_nearest_matrix_against_kdtree(oracle):
    rng = np.random.default_rng(3)
    coarse, fine = rng.random((300, 3)), rng.random((2000, 3))
    r, c, w = oracle.nearest_matrix(coarse, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(2000, 300))
    _, idx = cKDTree(coarse).query(fine)
    np.testing.assert_array_equal(NN.argmax(axis=1).A.ravel(), idx)
    np.testing.assert_allclose(NN.sum(axis=1).A.ravel(), 1.0)
    # exact ties share the weight equally (nearestMatrix.m:35-37)
    coarse = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]])
    fine = np.array([[0.5, 0, 0], [0.5, 0.5, 0], [0.1, 0.1, 0]])
    r, c, w = oracle.nearest_matrix(coarse, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(3, 3)).toarray()
    np.testing.assert_allclose(NN, [[0.5, 0.5, 0], [1 / 3, 1 / 3, 1 / 3], [1, 0, 0]])


def test_nearest_mode_root_plus_eight(oracle):
    """NEAREST: u = K(fine -> coarse) (NN (x) I3) f (Tools/Nearest/*.m)."""
    rng = np.random.default_rng(4)
    p, fine = cube_sphere(8), cube_sphere(32)
    r, c, w = oracle.nearest_matrix(p, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(len(fine), len(p)))
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=64, finePoints=fine, NN=NN)
    u = t.matvec(q, 0.01, 1.0, 6, 1)
    assert rel(u, oracle.direct(fine, NN @ q, p, 0.01)) < 1e-13


def test_reduced_pass_and_solve(oracle):
    rng = np.random.default_rng(5)
    p = rng.random((1200, 3))
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=30)
    u = t.reduced_apply(q, 0.02)
    ref = np.zeros_like(u)
    for leaf in np.unique(t.pointIndex):
        m = t.pointIndex == leaf
        ref[m] = oracle.direct(t.points[m], q[m], t.points[m], 0.02)
    assert rel(u, ref) < 1e-14
    f = t.reduced_solve(u, 0.02)
    assert rel(f, q) < 1e-8
    np.testing.assert_allclose(t.reduced_apply(f, 0.02), u, rtol=0, atol=1e-10 * np.abs(u).max())
