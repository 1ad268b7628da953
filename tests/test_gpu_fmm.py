"""HIP library against the oracle (GPU).

Tolerance: the FP64 form must reproduce the reference algorithm's
velocities within 1e-10 relative error (2-norm over all components).  The
per-level check-to-equivalent solves apply SVD factors instead of the
reference's per-node LU; both are backward stable, so they agree to far
below this bound.
"""
import ctypes
import os

import numpy as np
import pytest
import scipy.sparse as sp

from geometry import blobs, cube_sphere

pytestmark = pytest.mark.gpu
TOL = 1e-10
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available()
    return _t


@pytest.fixture(scope="module")
def selftest(nnk):
    nnk.lib()
    return ctypes.CDLL(os.path.join(os.path.dirname(nnk.library_path()), "libnnk_selftest.so"))


def _case(name):
    rng = np.random.default_rng(21)
    if name == "uniform":
        p = rng.random((6000, 3)); return p, p, 64, 1e-3, 1.0
    if name == "blobs":
        p = blobs(rng); return p, p, 16, 1e-3, 1.0
    if name == "distinct":
        p = rng.random((3000, 3)); return p, np.concatenate([rng.random((2000, 3)), p[:500]]), 40, 5e-3, 0.7
    if name == "sphere":
        p = cube_sphere(8); return p, p, 64, 0.01, 1.0
    if name == "shell":
        p = cube_sphere(20) + 0.02 * rng.random((2400, 3)); return p, p, 24, 2e-3, 1.0
    raise KeyError(name)


@pytest.mark.parametrize("name", ["uniform", "blobs", "distinct", "sphere", "shell"])
@pytest.mark.parametrize("res", [6, 4])
def test_matvec_matches_oracle(oracle, nnk, name, res):
    pot, tgt, cap, eps, mu = _case(name)
    q = np.random.default_rng(1).standard_normal((len(pot), 3))
    t = nnk.OcTree(pot, tgt, "nodeCapacity", cap)
    q = q[:int(t.potentials.sum())]
    fmm = nnk.KIFMM(t, [eps, mu], "coronaRes", res)
    u = fmm.computeVel(q)
    uo = oracle.Tree(pot, tgt, nodeCapacity=cap).matvec(q, eps, mu, res, 1)
    assert rel(u, uo) < TOL, rel(u, uo)


@pytest.mark.parametrize("precision", ["fp64", "fp64t"])
@pytest.mark.parametrize("name", ["blobs", "shell"])
def test_pair_layout_matches_class_tiles(oracle, nnk, name, precision):
    """The FP64 forms' per-offset M2L (the V-list pairs of one offset per
    job, each product into its own column, then a fixed-order sum per
    target), which the planner picks on sparse adaptive levels, against the
    class tiles on the same levels: equal to rounding, deterministic, and
    (FP64 form) equal to the oracle within the FP64 tolerance."""
    pot, tgt, cap, eps, mu = _case(name)
    t = nnk.OcTree(pot, tgt, "nodeCapacity", cap)
    q = np.random.default_rng(2).standard_normal((int(t.potentials.sum()), 3))
    f = nnk.KIFMM(t, [eps, mu], "coronaRes", 6, "precision", precision)
    lib, h = nnk.lib(), f._handle().h
    i32 = ctypes.POINTER(ctypes.c_int32)
    s0 = np.zeros(16, np.int32)
    d0 = np.zeros(16, np.int32)
    assert lib.nnk_fmm_launch_config(h, s0.ctypes.data_as(i32), d0.ctypes.data_as(i32), 16) == 0
    pl = np.flatnonzero(d0 == 1)
    print(f"[pairs] {name} {precision}: per-offset levels {pl.tolist()} of {np.flatnonzero(s0).tolist()}",
          flush=True)
    assert len(pl) >= 1 and (s0[pl] == 1).all()
    u1 = f.computeVel(q)
    np.testing.assert_array_equal(f.computeVel(q), u1)

    def setcfg(split, depth):
        s = np.zeros(16, np.int32)
        d = np.zeros(16, np.int32)
        s[pl], d[pl] = split, depth
        return lib.nnk_fmm_set_launch_config(h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16)

    assert setcfg(2, 1) != 0          # the per-offset jobs take no split-K
    assert setcfg(1, 0) == 0          # class tiles on the same levels
    u0 = f.computeVel(q)
    assert rel(u1, u0) < 1e-12, rel(u1, u0)   # summation order only, as for the autotuner's choices
    assert setcfg(1, 1) == 0
    np.testing.assert_array_equal(f.computeVel(q), u1)
    if precision == "fp64":
        uo = oracle.Tree(pot, tgt, nodeCapacity=cap).matvec(q, eps, mu, 6, 1)
        assert rel(u1, uo) < TOL, rel(u1, uo)


def test_golden_fixtures(nnk):
    for name in ("adaptive_res4.npz", "distinct_res6.npz"):
        g = dict(np.load(os.path.join(GOLD, name)))
        tg = g.get("targets", g["points"])
        t = nnk.OcTree(g["points"], tg, "nodeCapacity", int(g["cap"]))
        u = nnk.KIFMM(t, [float(g["eps"]), float(g["mu"])], "coronaRes", int(g["res"])).computeVel(g["q"])
        assert rel(u, g["u"]) < TOL, name
    g = dict(np.load(os.path.join(GOLD, "sphere_nearest.npz")))
    NN = sp.csc_matrix((g["nn_w"], (g["nn_row"], g["nn_col"])), shape=(len(g["fine"]), len(g["points"])))
    t = nnk.OcTree(g["points"], g["points"], "nodeCapacity", int(g["cap"]), "nearest", 1,
                   "finePoints", g["fine"], "NN", NN)
    f = nnk.KIFMM(t, [float(g["eps"]), float(g["mu"])])
    assert rel(f.computeVel(g["q"]), g["u"]) < TOL
    assert rel(f.reducedComputeVel(g["q"]), g["u_reduced"]) < TOL


def test_nearest_adaptive_matches_oracle(oracle, nnk):
    """NEAREST on a multi-level tree: two spheres, 16x fine points."""
    p = np.concatenate([cube_sphere(12), cube_sphere(12, centre=(2.5, 0.3, 0.1))])
    fine = np.concatenate([cube_sphere(48), cube_sphere(48, centre=(2.5, 0.3, 0.1))])
    r, c, w = oracle.nearest_matrix(p, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(len(fine), len(p)))
    q = np.random.default_rng(3).standard_normal(p.shape)
    t = nnk.OcTree(p, p, "nodeCapacity", 40, "nearest", 1, "finePoints", fine, "NN", NN)
    assert t.nodeLevel.max() >= 3
    u = nnk.KIFMM(t, [0.01, 1.0]).computeVel(q)
    uo = oracle.Tree(p, p, nodeCapacity=40, finePoints=fine, NN=NN).matvec(q, 0.01, 1.0, 6, 1)
    assert rel(u, uo) < TOL


def test_reduced_matches_oracle(oracle, nnk):
    pot, tgt, cap, eps, mu = _case("blobs")
    q = np.random.default_rng(2).standard_normal(pot.shape)
    t = nnk.OcTree(pot, tgt, "nodeCapacity", cap)
    u = nnk.KIFMM(t, [eps, mu]).reducedComputeVel(q)
    assert rel(u, oracle.Tree(pot, tgt, nodeCapacity=cap).reduced_apply(q, eps, mu)) < 1e-13


def test_direct_apply_matches_oracle(oracle, nnk):
    from importlib import import_module
    L = import_module("stokes-nnkifmm_amd._lib")
    rng = np.random.default_rng(4)
    s, q, t = rng.random((5000, 3)), rng.standard_normal((5000, 3)), rng.random((700, 3))
    u = np.empty_like(t)
    L.check(nnk.lib().nnk_direct_apply(L.ptr(s), 5000, L.ptr(q), L.ptr(t), 700, 1e-3, 1.2, L.ptr(u)))
    assert rel(u, oracle.direct(s, q, t, 1e-3, 1.2)) < 1e-13


def test_layouts_and_device_path(nnk, torch):
    pot, tgt, cap, eps, mu = _case("uniform")
    q = np.random.default_rng(5).standard_normal(pot.shape)
    t = nnk.OcTree(pot, tgt, "nodeCapacity", cap)
    u = nnk.KIFMM(t, [eps, mu]).computeVel(q)
    f1 = nnk.KIFMM(t, [eps, mu], "GMRES", 1, "format", 1)
    np.testing.assert_array_equal(f1.computeVel(q.ravel()), u.ravel())
    f2 = nnk.KIFMM(t, [eps, mu], "GMRES", 1, "format", 2)
    np.testing.assert_array_equal(f2.computeVel(q.T.ravel()), u.T.ravel())
    ud = nnk.KIFMM(t, [eps, mu]).computeVel(torch.tensor(q, device="cuda"))
    assert ud.is_cuda
    np.testing.assert_array_equal(ud.cpu().numpy(), u)


def test_linearity_at_scale(nnk, torch):
    """Size-independent property at 2^18 points: A(a x + b y) = a A x + b A y,
    and accuracy against the GPU direct sum on sampled targets."""
    from importlib import import_module
    L = import_module("stokes-nnkifmm_amd._lib")
    rng = np.random.default_rng(6)
    N = 1 << 18
    p = rng.random((N, 3))
    t = nnk.OcTree(p, p, "nodeCapacity", 64)
    f = nnk.KIFMM(t, [1e-3, 1.0])
    x = torch.tensor(rng.standard_normal((N, 3)), device="cuda")
    y = torch.tensor(rng.standard_normal((N, 3)), device="cuda")
    ax, ay, axy = f.computeVel(x), f.computeVel(y), f.computeVel(2.0 * x - 3.0 * y)
    # rounding through the ill-conditioned (cond ~ 2e10) solves sets the floor
    assert (torch.linalg.norm(axy - (2 * ax - 3 * ay)) / torch.linalg.norm(axy)).item() < 1e-11
    idx = np.sort(rng.choice(N, 1024, replace=False))
    tg = torch.tensor(p[idx], device="cuda")
    sd = torch.tensor(p, device="cuda")
    ud = torch.empty((1024, 3), dtype=torch.float64, device="cuda")
    L.check(nnk.lib().nnk_direct_apply_device(ctypes.c_void_p(sd.data_ptr()), N, ctypes.c_void_p(x.data_ptr()),
                                              ctypes.c_void_p(tg.data_ptr()), 1024, 1e-3, 1.0,
                                              ctypes.c_void_p(ud.data_ptr()), None))
    torch.cuda.synchronize()
    assert (torch.linalg.norm(ax[idx] - ud) / torch.linalg.norm(ud)).item() < 1e-4


@pytest.mark.parametrize("bn", [64, 32])
def test_gemm_selftest(selftest, bn):
    """The FP64 gather GEMM at both tile widths the library launches (64
    columns for M2L, 32 for the expansions, M2M, L2L and solves) against the
    host product."""
    d = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    rng = np.random.default_rng(7)
    M = K = 456; Mpad = 480; Kpad = 464; nb, nx, N, nc = 3, 50, 101, 120
    At = np.zeros((nb, Kpad, Mpad)); At[:, :K, :M] = rng.standard_normal((nb, K, M))
    X = rng.standard_normal((nx, K))
    g = rng.integers(-1, nx, (nb, N)).astype(np.int32)
    ci = rng.permutation(nc)[:N].astype(np.int32)
    C0 = rng.standard_normal((nc, K))
    ref = C0.copy()
    for n in range(N):
        acc = sum(At[b, :K, :M].T @ X[g[b, n]] for b in range(nb) if g[b, n] >= 0)
        ref[ci[n]] = acc + C0[ci[n]]
    C = C0.copy()
    assert selftest.nnk_selftest_gemm(M, K, Mpad, Kpad, At.ctypes.data_as(d), nb, X.ctypes.data_as(d), nx, K,
                                      g.ctypes.data_as(ip), ci.ctypes.data_as(ip), N, C.ctypes.data_as(d),
                                      nc, K, 1, bn) == 0
    np.testing.assert_allclose(C, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
