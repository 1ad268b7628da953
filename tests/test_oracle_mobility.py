"""Mobility-layer oracle pins (CPU).

  * Stokes drag: the resistance problem S f = e_x on a unit sphere gives a
    total force 6 pi mu a (the reference's own check,
    Data Generation/DirectMethodComparison.m:32-34), here with the NEAREST
    quadrature of the one-sphere configuration;
  * the matrix-free B / B^T / BT / BT^T are mutual transposes;
  * oracle GMRES against numpy.linalg.solve on a small dense system.
"""
import numpy as np
import scipy.sparse as sp

from geometry import cube_sphere


def _sphere(oracle, n=8, nf=32):
    p, fine = cube_sphere(n), cube_sphere(nf)
    r, c, w = oracle.nearest_matrix(p, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(len(fine), len(p)))
    return p, fine, NN


def test_transposes(oracle):
    from oracle.mobility import MobilityOracle
    p, fine, NN = _sphere(oracle, 4, 8)
    t = oracle.Tree(p, p, nodeCapacity=64, finePoints=fine, NN=NN)
    pm = NN.T @ fine
    m = MobilityOracle(t, 0.05, 1.0, [0, len(p)], p, np.asarray(NN.sum(0)).ravel(), pm)
    rng = np.random.default_rng(0)
    f, y = rng.standard_normal(3 * len(p)), rng.standard_normal(6)
    assert abs(m.B(f) @ y - f @ m.Bt(y)) < 1e-12 * np.abs(f).sum()
    v = rng.standard_normal(3 * len(p))
    assert abs(m.BT(y) @ v - y @ m.BTt(v)) < 1e-12 * np.abs(v).sum()
    np.testing.assert_allclose(m.A1, m.A1.T, atol=1e-12)


def test_sphere_drag_6pi(oracle):
    from oracle.mobility import gmres
    p, fine, NN = _sphere(oracle)
    t = oracle.Tree(p, p, nodeCapacity=64, finePoints=fine, NN=NN)
    vel = np.zeros((len(p), 3))
    vel[:, 0] = 1.0
    op = lambda f: t.matvec(f.reshape(3, -1).T, 0.01, 1.0, 6, 1).T.ravel()
    f, info = gmres(op, vel.T.ravel(), 1e-8, 400)
    assert info["flag"] == 0
    F = (NN @ f.reshape(3, -1).T).sum(0)     # force carried by the fine points
    assert abs(F[0] / (6 * np.pi) - 1) < 0.02 and np.abs(F[1:]).max() < 1e-6, F


def test_gmres_dense(oracle):
    from oracle.mobility import gmres
    rng = np.random.default_rng(1)
    A = rng.standard_normal((60, 60)) + 8 * np.eye(60)
    b = rng.standard_normal(60)
    x, info = gmres(lambda v: A @ v, b, 1e-12, 60)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9)
