"""Mobility-layer oracle pins (CPU).

  * Stokes drag: the resistance problem S f = e_x on a unit sphere gives a
    total force 6 pi mu a (the reference's own check,
    Data Generation/DirectMethodComparison.m:32-34), here with the NEAREST
    quadrature of the one-sphere configuration;
  * the matrix-free B / B^T / BT / BT^T are mutual transposes;
  * oracle GMRES against numpy.linalg.solve on a small dense system;
  * the preconditioner and the saddle-point operator against dense
    matrices written from genMatrixComponets.m, ReducedSolve.m and
    MobilityPrecon.m on the one-sphere tree.
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


def test_precon_matches_dense_restatement(oracle):
    """MobilityOracle.precon against MobilityPrecon.m:17-39 evaluated with
    dense matrices on the one-sphere tree (root + 8 mutually adjacent
    leaves, so the coronaRes-4 FMM equals the direct sum):
      S4 = K(fine -> points) (NN kron I3)                      (AoS)
      R  = blkdiag_L K(fine_L -> points_L) (NN[fine_L, L] kron I3)
           (ReducedSolve.m:93-111: fine_L = fine points with NN weight in
           leaf L; only leaf L's columns of NN)
      B  = [kron(I3, af); AM], BT = [kron(I3, au), AOm], A1 = B B^T,
      A2 = BT^T BT (genMatrixComponets.m:34-100, NEAREST branch)
      b_f = R^-1 f - R^-1 BT out,  out = -A2 \\ (BT^T S4 B^T (A1 \\ UOm))
    A sign or transpose slip in the oracle changes the result by O(1)."""
    from oracle.mobility import MobilityOracle
    eps, mu = 0.01, 1.0
    p, fine, NN = _sphere(oracle)
    N, Q = len(p), len(fine)
    t = oracle.Tree(p, p, nodeCapacity=64, finePoints=fine, NN=NN)
    assert t.nodeCount == 9
    af = np.asarray(NN.sum(0)).ravel()
    pm = NN.T @ fine
    m = MobilityOracle(t, eps, mu, [0, N], p, af, pm)

    I3 = sp.identity(3, format="csr")
    NNk = sp.kron(NN, I3).tocsc()                         # (3Q, 3N) AoS
    K = oracle.kernel(fine, p, eps, mu)                   # (3N, 3Q) AoS, rows = targets
    S4_aos = np.asarray((NNk.T @ K.T).T)
    perm = (np.arange(3)[:, None] + 3 * np.arange(N)[None, :]).ravel()   # SoA slot -> AoS slot
    S4 = S4_aos[np.ix_(perm, perm)]
    R_aos = np.zeros((3 * N, 3 * N))
    leaf = t.pointIndex[t.potentials]
    NNc = NN.tocsc()
    for L in np.unique(leaf):
        cols = np.where(leaf == L)[0]
        sub = NNc[:, cols]
        rows = np.where(np.asarray(abs(sub).sum(1)).ravel() > 0)[0]
        blk = sp.kron(sub[rows], I3).toarray()
        Kl = oracle.kernel(fine[rows], p[cols], eps, mu)
        ia = (3 * cols[:, None] + np.arange(3)[None, :]).ravel()
        R_aos[np.ix_(ia, ia)] = Kl @ blk
    R = R_aos[np.ix_(perm, perm)]
    Z = np.zeros((1, N))
    x1, x2, x3 = (pm[:, d][None, :] for d in range(3))
    B = np.vstack([np.kron(np.eye(3), af[None, :]),
                   np.block([[Z, -x3, x2], [x3, Z, -x1], [-x2, x1, Z]])])
    z = np.zeros((N, 1))
    r1, r2, r3 = (p[:, d][:, None] for d in range(3))
    BT = np.hstack([np.kron(np.eye(3), -np.ones((N, 1))),
                    np.block([[z, -r3, r2], [r3, z, -r1], [-r2, r1, z]])])
    A1, A2 = B @ B.T, BT.T @ BT
    np.testing.assert_allclose(m.A1, A1, rtol=1e-12, atol=1e-12 * np.abs(A1).max())
    np.testing.assert_allclose(m.A2, A2, rtol=1e-12, atol=1e-12 * np.abs(A2).max())

    rng = np.random.default_rng(3)
    x = rng.standard_normal(3 * N + 6)
    f, uo = x[:3 * N], x[3 * N:]
    out = -np.linalg.solve(A2, BT.T @ (S4 @ (B.T @ np.linalg.solve(A1, uo))))
    bf = np.linalg.solve(R, f) - np.linalg.solve(R, BT @ out)
    want = np.concatenate([bf, out])
    got = m.precon(x)
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want), np.linalg.norm(got - want) / np.linalg.norm(want)
    # the operator and B/BT of the saddle-point matvec (MobilityMatrixMulti.m:31-35) the same way
    # m.apply uses the res-6 operator: on root + 8 it is the direct sum as well
    y = m.apply(x)
    want_a = np.concatenate([S4 @ f + BT @ uo, B @ f])
    assert np.linalg.norm(y - want_a) <= 1e-10 * np.linalg.norm(want_a)
