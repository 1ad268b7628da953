"""The benchmark configurations at (or near) their stated sizes (GPU).

  * the headline mixed forms of the FMM matvec (two-part FP16 at coronaRes 9
    with the pseudo-inverses cut at 3e-7, three-part BF16 at coronaRes 8) on
    131,072 uniform points (32 per leaf, as at 2^20) against the oracle's
    FP64 direct sum: below 1e-6;
  * a reduced spheroid suspension above a wall (8 prolate spheroids, aspect
    3, eps 0.005, NEAREST quadrature): mobilityMatrixMulti and
    mobilityPrecon against the oracle (1e-10 / 1e-9), and the refined mixed
    solve against the FP64 solve at the configuration's tolerance 1e-6;
  * batched single-body problems of 200 and 5,000 force points (4x refined
    quadrature: up to 80,736 sources) against the oracle's direct sum;
  * the squirmer pair at its stated size (1,536 force / 24,576 quadrature
    points per body, eps 0.01) for B2 in {-5, 0, 5}: adjoint gradients
    against central finite differences with h = 1e-6.
"""
import importlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _geo():
    return importlib.import_module("stokes-nnkifmm_amd.geometries")


@pytest.mark.parametrize("precision,res,svd_tol", [("mixedh", 9, 3e-7), ("mixed3", 8, 1e-8)])
def test_headline_forms_below_1e6(oracle, nnk, precision, res, svd_tol):
    rng = np.random.Generator(np.random.PCG64(2024))
    p = rng.random((1 << 17, 3))
    q = rng.standard_normal(p.shape)
    tree = nnk.OcTree(p, p, "nodeCapacity", 64)
    f = nnk.KIFMM(tree, [1e-3, 1.0], "coronaRes", res, "precision", precision, "svdTol", svd_tol)
    u = f.computeVel(q)
    idx = np.sort(rng.choice(len(p), 2000, replace=False))
    ud = oracle.direct(p, q, p[idx], 1e-3)
    err = rel(u[idx], ud)
    print(f"{precision} res {res} svdTol {svd_tol}: rel err vs FP64 direct sum {err:.3e}")
    assert err < 1e-6, err


@pytest.fixture(scope="module")
def suspension(nnk, oracle):
    """8 spheroids (384 force / 6,144 quadrature points each) in an 8^3 box
    above a 12 x 12 wall (48 x 48 quadrature), eps 0.005, plus the oracle's
    twin of the same system."""
    import scipy.sparse as sp
    from oracle.mobility import MobilityOracle
    sw = _geo().spheroid_suspension(nbody=8, n=8, box=8.0, wall_n=12, seed=3)
    sw.genTree("nodeCapacity", 64, "nearest", 1)
    sw.genKIFMM([0.005, 1.0], "GMRES", 1, "format", 2)
    B, BT, A1, A2 = nnk.genMatrixComponets(sw)
    pts, fpts, NNall = sw.getSwimmerBnd()
    t = oracle.Tree(pts, pts, nodeCapacity=64, finePoints=fpts, NN=NNall)
    _, _, NNs, rot, frot = sw.getSwimmers()
    sizes, _ = sw.getSwimmerSizes()
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    w = np.asarray(sp.csc_matrix(NNs).sum(0)).ravel()
    pm = sp.csc_matrix(NNs).T @ frot
    m = MobilityOracle(t, 0.005, 1.0, ptr, rot, w, pm)
    return sw, B, BT, A1, A2, m


def test_suspension_operators_match_oracle(nnk, suspension):
    sw, B, BT, A1, A2, m = suspension
    assert sw.Tree.nodeCount > 100 and max(sw.Tree.nodeLevel) >= 4   # an adaptive multi-level tree
    assert abs(A1 - m.A1).max() < 1e-10 * abs(m.A1).max()
    assert abs(A2 - m.A2).max() < 1e-10 * abs(m.A2).max()
    x = np.random.default_rng(4).standard_normal(m.n)
    assert rel(nnk.mobilityMatrixMulti(x, sw, B, BT), m.apply(x)) < 1e-10
    assert rel(nnk.mobilityPrecon(x, sw, B, BT, A1, A2, 1e-6, 100), m.precon(x)) < 1e-9


def test_suspension_refined_solve(nnk, suspension):
    """FP64 GMRES (preconditioned) and iterative refinement around the mixed
    form, both to the configuration's 1e-6: the refined answer meets 1e-6 on
    the FP64 system and its body velocities agree with the FP64 solve's."""
    sw, B, BT, A1, A2, m = suspension
    mob = importlib.import_module("stokes-nnkifmm_amd.mobility")
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    x64 = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-6, 400, True)
    i64 = dict(mob.last_solve)
    xr = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-6, 400, True, innerTol=1e-3)
    ir = dict(mob.last_solve)
    assert i64["flag"] == 0 and ir["flag"] == 0, (i64, ir)
    r = m.apply(xr) - b                       # residual under the oracle's FP64 operator
    assert np.linalg.norm(r) <= 1.05e-6 * np.linalg.norm(b)
    nb = 6 * sw.swimmerNo()
    assert rel(xr[-nb:], x64[-nb:]) < 1e-4


def test_suspension_fgmres_solve(nnk, suspension):
    """Flexible GMRES (FP64 Arnoldi on the FP64 operator, every basis vector
    preconditioned by a mixed-form inner GMRES to 1e-2): meets 1e-6 on the
    oracle's FP64 system in a few outer steps, and its body velocities agree
    with the FP64 solve's."""
    sw, B, BT, A1, A2, m = suspension
    mob = importlib.import_module("stokes-nnkifmm_amd.mobility")
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    x64 = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-6, 400, True)
    xf = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-6, 400, True, innerTol=1e-2, outer="fgmres")
    info = dict(mob.last_solve)
    print(f"[fgmres] {info}", flush=True)
    assert info["flag"] == 0 and 1 <= info["outer"] <= 10, info
    r = m.apply(xf) - b
    assert np.linalg.norm(r) <= 1.05e-6 * np.linalg.norm(b)
    nb = 6 * sw.swimmerNo()
    assert rel(xf[-nb:], x64[-nb:]) < 1e-4
    # inner solves on the coarser coronaRes-4 operator: the outer FP64
    # Arnoldi still decides convergence, so the answer meets the same bar
    x4 = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-6, 400, True, innerTol=1e-2, outer="fgmres", innerRes=4)
    info = dict(mob.last_solve)
    print(f"[fgmres, res-4 inner] {info}", flush=True)
    assert info["flag"] == 0 and 1 <= info["outer"] <= 20, info
    r = m.apply(x4) - b
    assert np.linalg.norm(r) <= 1.05e-6 * np.linalg.norm(b)
    assert rel(x4[-nb:], x64[-nb:]) < 1e-4


def _oracle_apply(oracle, sw, eps, x):
    pout, fout, NN, prot, frot = sw.getSwimmers()
    n = len(pout)
    f = x[:3 * n].reshape(3, n).T
    U, W = x[3 * n:3 * n + 3], x[3 * n + 3:3 * n + 6]
    u = oracle.direct(fout, NN @ f, pout, eps) - U - np.cross(W, prot)
    w = np.asarray(NN.sum(axis=0)).ravel()
    pm = NN.T @ frot
    F = (w[:, None] * f).sum(0)
    M = np.cross(pm, f).sum(0)
    return np.concatenate([u.T.ravel(), F, M])


def test_batched_problems_at_stated_sizes(oracle, nnk):
    """Two 5,000-point and two 200-point problems (the ends of the stated
    range), eps in [0.002, 0.02]."""
    geo = _geo()
    big, eb = geo.batch_bodies(nprob=2, nmin=5000, nmax=5000, seed=21)
    small, es = geo.batch_bodies(nprob=2, nmin=200, nmax=200, seed=22)
    bodies, eps = big + small, np.concatenate([eb, es])
    sizes = [len(b.Geometries[0]) for b in bodies]
    assert max(sizes) >= 5000 and min(sizes) <= 220, sizes
    bm = nnk.BatchedMobility(bodies, eps, hipGraph=True)
    x = np.random.default_rng(0).standard_normal(bm.n)
    for p, (sw, xp, yp) in enumerate(zip(bodies, bm.split(x), bm.split(bm.apply(x)))):
        ref = _oracle_apply(oracle, sw, eps[p], xp)
        assert np.linalg.norm(yp - ref) <= 1e-12 * np.linalg.norm(ref), p
    b = bm.rhs()
    z = bm.solve(b, 1e-6, 300)
    assert (bm.last_solve["flag"] == 0).all(), bm.last_solve
    for p, (sw, zp, bp) in enumerate(zip(bodies, bm.split(z), bm.split(b))):
        r = _oracle_apply(oracle, sw, eps[p], zp) - bp
        assert np.linalg.norm(r) <= 1.05e-6 * np.linalg.norm(bp), p


@pytest.mark.parametrize("B2", [-5.0, 0.0, 5.0])
def test_squirmer_pair_gradients_full_size(nnk, B2):
    """Gradients of U_x (body 0) and U_z (body 1) with respect to every
    force-point position, every quadrature-point position and eps, from one
    transpose GMRES, against central differences (h = 1e-6) of forward
    solves along random directions that move every node at once (the
    nearest map and the slip held fixed), and along eps."""
    swm = importlib.import_module("stokes-nnkifmm_amd.swimmers")
    geo = _geo()
    eps, h, tol = 0.01, 1e-6, 1e-13

    def system(dx=None, dy=None):
        sw = geo.squirmer_pair(n=16, refine=4, gap=0.5, B1=1.0, B2=B2)
        for b in range(2):
            Bm = swm._basis(*sw.b[b])            # world = Bm body + x0
            if dx is not None:
                sw.Geometries[b] = sw.Geometries[b] + dx[b] @ Bm
            if dy is not None:
                sw.GeometriesFine[b] = sw.GeometriesFine[b] + dy[b] @ Bm
        return sw

    base = nnk.BatchedMobility([system()], [eps])
    N, Q = int(base.npts[0]), int(base.nfine[0])
    assert N == 2 * 1536 and Q == 2 * 24576
    b = base.rhs()
    z = base.solve(b, tol, 600)
    assert base.last_solve["flag"][0] == 0
    n = len(z)
    slots = [3 * N + 0 * 2 + 0, 3 * N + 2 * 2 + 1]
    both = nnk.BatchedMobility([system(), system()], [eps, eps])
    e = np.zeros(2 * n)
    e[slots[0]] = 1.0
    e[n + slots[1]] = 1.0
    lam = both.solve(e, tol, 600, transpose=True)
    assert (both.last_solve["flag"] == 0).all()
    gx, gy, ge = both.gradient(np.concatenate([z, z]), lam)

    rng = np.random.default_rng(int(B2) + 10)
    vx = rng.standard_normal((2, N // 2, 3))
    vy = rng.standard_normal((2, Q // 2, 3))

    def J(sw, e_):
        return nnk.BatchedMobility([sw], [e_]).solve(b, tol, 600)[slots]

    fd = {"eps": (J(system(), eps + h) - J(system(), eps - h)) / (2 * h),
          "x": (J(system(dx=h * vx), eps) - J(system(dx=-h * vx), eps)) / (2 * h),
          "y": (J(system(dy=h * vy), eps) - J(system(dy=-h * vy), eps)) / (2 * h)}
    adj = {"eps": ge, "x": np.array([(g * vx.reshape(-1, 3)).sum() for g in gx]),
           "y": np.array([(g * vy.reshape(-1, 3)).sum() for g in gy])}
    for k in fd:
        print(f"B2={B2} d/d{k}: fd {fd[k]}, adjoint {adj[k]}")
        assert np.abs(adj[k]).max() > 0
        assert np.abs(fd[k] - adj[k]).max() <= 2e-5 * np.abs(adj[k]).max(), (k, fd[k], adj[k])
