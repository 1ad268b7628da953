"""Mobility path on the GPU against the oracle (GPU).

nearestMatrix is compared bit-for-bit; reducedSolve, mobilityMatrixMulti
and mobilityPrecon to 1e-10 (reducedSolve: LU on both sides); the GMRES
solves to their tolerance; the one-sphere mobility problem against Stokes'
law U = F / (6 pi mu a).
"""
import numpy as np
import pytest
import scipy.sparse as sp

from geometry import cube_sphere

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _nearest_cases():
    rng = np.random.default_rng(0)
    yield cube_sphere(8), cube_sphere(32)
    yield rng.random((500, 3)), rng.random((4000, 3))
    yield np.array([[0.0, 0, 0], [1, 0, 0], [0, 1, 0]]), np.array([[.5, 0, 0], [.5, .5, 0]])
    g = np.stack(np.meshgrid(*[np.arange(8.0)] * 3, indexing="ij"), -1).reshape(-1, 3)
    yield g, g + 0.5                                           # 8-way ties at cell centres
    yield g, rng.integers(-2, 17, (3000, 3)) / 2.0             # lattice ties, points outside the grid
    yield np.c_[rng.random((2000, 2)), np.zeros(2000)], rng.random((3000, 3)) * 3 - 1   # flat coarse set
    yield np.zeros((5, 3)), rng.random((50, 3))                 # coincident coarse points
    s = cube_sphere(20)
    yield s, np.r_[np.zeros((1, 3)), 1.3 * s[:500], 0.2 * rng.standard_normal((500, 3))]
    yield rng.random((3000, 3)) * 1e6 + 1e7, rng.random((2000, 3)) * 1e6 + 1e7
    yield cube_sphere(32, 1.0, (0, 0, 0)), cube_sphere(128)    # 6144 x 98304


def test_nearest_matrix_bit_exact(oracle, nnk):
    for coarse, fine in _nearest_cases():
        NN = nnk.nearestMatrix(coarse, fine)
        r, c, w = oracle.nearest_matrix(coarse, fine)
        ref = sp.csc_matrix((w, (r, c)), shape=(len(fine), len(coarse)))
        ref.sort_indices()
        np.testing.assert_array_equal(NN.indptr, ref.indptr)
        np.testing.assert_array_equal(NN.indices, ref.indices)
        np.testing.assert_array_equal(NN.data, ref.data)


def test_reduced_solve_matches_oracle(oracle, nnk):
    rng = np.random.default_rng(1)
    p = rng.random((2000, 3))
    v = rng.standard_normal(p.shape)
    t = nnk.OcTree(p, p, "nodeCapacity", 40)
    f = nnk.KIFMM(t, [0.02, 1.0]).reducedSolve(v)
    fo = oracle.Tree(p, p, nodeCapacity=40).reduced_solve(v, 0.02, 1.0)
    assert rel(f, fo) < 1e-10
    # NEAREST blocks (K(fine -> leaf) kron(NN, I3))
    c, fine = cube_sphere(8), cube_sphere(32)
    NN = nnk.nearestMatrix(c, fine)
    v = rng.standard_normal(c.shape)
    t = nnk.OcTree(c, c, "nodeCapacity", 64, "nearest", 1, "finePoints", fine, "NN", NN)
    f = nnk.KIFMM(t, [0.01, 1.0]).reducedSolve(v)
    fo = oracle.Tree(c, c, nodeCapacity=64, finePoints=fine, NN=NN).reduced_solve(v, 0.01, 1.0)
    assert rel(f, fo) < 1e-10


def test_reduced_solve_large_leaf_blocks(oracle, nnk):
    """Eight leaves of about 375 points (blocks of about 1,100 unknowns:
    many 32-row blocks of the inverse with a partial last one, several
    64-column workgroups per leaf, and several 256-row workgroups per leaf
    in the apply) against the oracle's per-leaf LU solve."""
    rng = np.random.default_rng(11)
    p = rng.random((3000, 3))
    v = rng.standard_normal(p.shape)
    t = nnk.OcTree(p, p, "nodeCapacity", 500)
    f = nnk.KIFMM(t, [0.05, 1.0]).reducedSolve(v)
    fo = oracle.Tree(p, p, nodeCapacity=500).reduced_solve(v, 0.05, 1.0)
    print(f"[reduced] large blocks: rel diff {rel(f, fo):.3e}", flush=True)
    assert rel(f, fo) < 1e-9


def _system(nnk, oracle, two_bodies):
    """One sphere (NEAREST, 384 / 6144 points) or two spheres over a plane
    wall (multi-level tree)."""
    sw = nnk.Swimmers()
    c, fine = cube_sphere(8), cube_sphere(32)
    NN = nnk.nearestMatrix(c, fine)
    rng = np.random.default_rng(2)
    centres = [[0.0, 0.0, 0.0]] if not two_bodies else [[-1.4, 0.2, 0.6], [1.3, -0.1, 0.8]]
    for i, x0 in enumerate(centres):
        ang = 0.3 * (i + 1)
        b1 = [np.cos(ang), np.sin(ang), 0.0]
        b2 = [-np.sin(ang), np.cos(ang), 0.0]
        sw.addSwimmer(c, fine, NN, x0, b1 + b2, np.zeros_like(c), [6 * np.pi, 0.3 * i, 0.0], [0.0, 0.0, 0.2 * i])
    if two_bodies:
        g = np.linspace(-3, 3, 16)
        X, Y = np.meshgrid(g, g)
        wall = np.stack([X.ravel(), Y.ravel(), -0.8 * np.ones(X.size)], 1)
        wf = np.stack([np.repeat(X.ravel(), 1), Y.ravel(), -0.8 * np.ones(X.size)], 1)
        sw.updateBnd(wall, wf, sp.identity(len(wall), format="csc"), np.zeros_like(wall))
    sw.genTree("nodeCapacity", 64 if not two_bodies else 48, "nearest", 1)
    sw.genKIFMM([0.01, 1.0], "GMRES", 1, "format", 2)
    B, BT, A1, A2 = nnk.genMatrixComponets(sw)
    # oracle twin built from the same geometry
    from oracle.mobility import MobilityOracle
    pts, fpts, NNall = sw.getSwimmerBnd()
    t = oracle.Tree(pts, pts, nodeCapacity=sw.Tree.arguments["nodeCapacity"], finePoints=fpts, NN=NNall)
    _, _, NNs, rot, frot = sw.getSwimmers()
    sizes, _ = sw.getSwimmerSizes()
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    w = np.asarray(sp.csc_matrix(NNs).sum(0)).ravel()
    pm = sp.csc_matrix(NNs).T @ frot
    m = MobilityOracle(t, 0.01, 1.0, ptr, rot, w, pm)
    return sw, B, BT, A1, A2, m


@pytest.mark.parametrize("two_bodies", [False, True])
def test_operators_match_oracle(oracle, nnk, two_bodies):
    sw, B, BT, A1, A2, m = _system(nnk, oracle, two_bodies)
    assert abs(A1 - m.A1).max() < 1e-10 * abs(m.A1).max()
    assert abs(A2 - m.A2).max() < 1e-10 * abs(m.A2).max()
    x = np.random.default_rng(3).standard_normal(m.n)
    assert rel(nnk.mobilityMatrixMulti(x, sw, B, BT), m.apply(x)) < 1e-10
    assert rel(nnk.mobilityPrecon(x, sw, B, BT, A1, A2, 1e-6, 100), m.precon(x)) < 1e-9


@pytest.mark.parametrize("precon", [False, True])
def test_sphere_mobility_stokes_law(oracle, nnk, precon):
    sw = nnk.Swimmers()
    c, fine = cube_sphere(8), cube_sphere(32)
    NN = nnk.nearestMatrix(c, fine)
    sw.addSwimmer(c, fine, NN, [0, 0, 0], [1, 0, 0, 0, 1, 0], np.zeros_like(c), [6 * np.pi, 0, 0], [0, 0, 0])
    out = nnk.mobilityProblem(sw, lambda t, x: np.zeros_like(x), 0.0, np.r_[0, 0, 0, 1, 0, 0, 0, 1, 0],
                              [0.01, 1.0], ["nodeCapacity", 64, "nearest", 1], 0, 1, 0.1, 1e-6, 200,
                              [1e-6, 100] if precon else [])
    info = nnk.mobility.last_solve
    assert info["flag"] == 0 and info["relres"] <= 1e-6
    U = out[0:3]
    assert abs(U[0] - 1) < 0.02 and np.abs(U[1:]).max() < 1e-6, U
    assert np.abs(out[3:]).max() < 1e-6   # no rotation


@pytest.mark.parametrize("precon", [False, True])
def test_gmres_matches_oracle(oracle, nnk, precon):
    sw, B, BT, A1, A2, m = _system(nnk, oracle, True)
    from importlib import import_module
    mob = import_module("stokes-nnkifmm_amd.mobility")
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    x = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-10, 300, precon)
    info = dict(mob.last_solve)
    xo, infoo = m.gmres(b, 1e-10, 300, precon)
    assert info["flag"] == 0 and infoo["flag"] == 0
    assert abs(info["iter"] - infoo["iter"]) <= 2, (info["iter"], infoo["iter"])
    assert rel(x, xo) < 1e-7


def test_fgmres_outer_iteration(oracle, nnk):
    """Flexible GMRES on the two-sphere + wall system.  With an FP64 inner
    operator solved to 1e-12, the first preconditioned direction already
    solves the system: one outer step (the flexible Arnoldi relation
    A Z = V H).  With a mixed inner operator at 1e-2 it takes a few steps
    and meets 1e-8 under the oracle's operator, matching the oracle's
    GMRES solution."""
    sw, B, BT, A1, A2, m = _system(nnk, oracle, True)
    from importlib import import_module
    mob = import_module("stokes-nnkifmm_amd.mobility")
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    x1 = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-8, 300, True, innerTol=1e-12, innerPrecision="fp64",
                              outer="fgmres")
    i1 = dict(mob.last_solve)
    assert i1["flag"] == 0 and i1["outer"] == 1, i1
    x2 = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-8, 300, True, innerTol=1e-2, outer="fgmres")
    i2 = dict(mob.last_solve)
    assert i2["flag"] == 0 and 2 <= i2["outer"] <= 12, i2
    for x in (x1, x2):
        assert np.linalg.norm(m.apply(x) - b) <= 1.05e-8 * np.linalg.norm(b)
    xo, _ = m.gmres(b, 1e-10, 300, True)
    assert rel(x2, xo) < 1e-6


@pytest.mark.parametrize("offset", [0.0, 0.3])
def test_sphere_gravity_sedimentation_and_offset_torque(nnk, offset):
    """mobilityProblemGravity (mobilityProblemGravity.m:25-74): F = G and
    M = -(4/3) pi offset (b3 x G) (Swimmers.m:372-385).  A sphere sediments
    at U = G / (6 pi mu a) and turns at Omega = M / (8 pi mu a^3) (Stokes'
    laws, to the discretisation's 2 %); with b1 = e_x, b2 = e_z the body
    axis b3 = -e_y, so G = -g e_z gives M = -(4/3) pi offset g e_x."""
    sw = nnk.Swimmers()
    c, fine = cube_sphere(8), cube_sphere(32)
    NN = nnk.nearestMatrix(c, fine)
    sw.addSwimmer(c, fine, NN, [0, 0, 0], [1, 0, 0, 0, 0, 1], np.zeros_like(c), [0, 0, 0], [0, 0, 0])
    g = 6 * np.pi
    out = nnk.mobilityProblemGravity(sw, lambda t, x: np.zeros_like(x), 0.0, np.r_[0, 0, 0, 1, 0, 0, 0, 0, 1],
                                     [0.01, 1.0], ["nodeCapacity", 64, "nearest", 1], 0, 1, 0.1, 1e-8, 300,
                                     [1e-6, 100], [0.0, 0.0, -g, offset])
    info = nnk.mobility.last_solve
    assert info["flag"] == 0
    np.testing.assert_allclose(sw.F[0], [0, 0, -g])
    M = -4.0 / 3.0 * np.pi * offset * np.cross([0, -1, 0], [0, 0, -g])
    np.testing.assert_allclose(sw.Moment[0], M, atol=1e-14)
    U = out[0:3]
    assert abs(U[2] + 1) < 0.02 and np.abs(U[:2]).max() < 1e-6, U
    # out[3:6] = Omega x b1, out[6:9] = Omega x b2 with b1 = e_x, b2 = e_z
    w1, w2 = out[3:6], out[6:9]
    Om = np.array([-w2[1], w2[0], w1[1]])
    np.testing.assert_allclose(np.cross(Om, [1, 0, 0]), w1, atol=1e-12)
    want = M / (8 * np.pi)
    if offset == 0.0:
        assert np.abs(Om).max() < 1e-6, Om
    else:
        assert abs(Om[0] / want[0] - 1) < 0.03 and np.abs(Om[1:]).max() < 1e-6 * abs(want[0]) + 1e-9, (Om, want)
