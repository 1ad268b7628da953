"""Mixed-precision form (GPU): truncated-SVD pseudo-inverses, split-BF16
M2L, FP32-math direct sums.

Its numbers differ from the reference's by design, so it is checked against
the FP64 direct sum (the accuracy the form promises) and against the FP64
form of this library (the two share everything but precision and the
pseudo-inverse cut).  Tolerances at 6 points per corona edge, where the FP64
form itself is at 4e-6 to 2.2e-5 against the direct sum: the mixed form
stays below 3e-5 and adds at most 5e-6 to the FP64 form's error.
"""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from geometry import blobs, cube_sphere

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("case", ["uniform", "blobs", "distinct"])
def test_mixed_against_direct_and_fp64(oracle, nnk, case):
    rng = np.random.default_rng(31)
    if case == "uniform":
        p = t = rng.random((20000, 3)); cap, eps = 64, 1e-3
    elif case == "blobs":
        p = t = blobs(rng); cap, eps = 16, 1e-3
    else:
        p = rng.random((8000, 3)); t = np.concatenate([rng.random((3000, 3)), p[:1000]]); cap, eps = 40, 5e-3
    tree = nnk.OcTree(p, t, "nodeCapacity", cap)
    q = rng.standard_normal((int(tree.potentials.sum()), 3))
    u64 = nnk.KIFMM(tree, [eps, 1.0]).computeVel(q)
    um = nnk.KIFMM(tree, [eps, 1.0], "precision", "mixed").computeVel(q)
    ud = oracle.direct(tree.getPotentialPoints(), q, tree.getTargetPoints(), eps)
    e64, em = rel(u64, ud), rel(um, ud)
    assert em < 3e-5 and em < e64 + 5e-6, (e64, em)
    assert rel(um, u64) < 5e-5


@pytest.mark.parametrize("precision", ["mixed3", "mixedh"])
def test_below_1e6_at_res8(oracle, nnk, precision):
    """The three-part BF16 and the two-part FP16 forms at 8 points per corona
    edge stay below 1e-6 against the FP64 direct sum (the two-part BF16
    form stops near 3e-6)."""
    rng = np.random.default_rng(8)
    p = rng.random((30000, 3))
    q = rng.standard_normal(p.shape)
    tree = nnk.OcTree(p, p, "nodeCapacity", 64)
    u3 = nnk.KIFMM(tree, [1e-3, 1.0], "coronaRes", 8, "precision", precision).computeVel(q)
    idx = rng.choice(len(p), 2000, replace=False)
    ud = oracle.direct(p, q, p[idx], 1e-3)
    assert rel(u3[idx], ud) < 1e-6


def test_mixedh_scaling_is_input_invariant(nnk):
    """The FP16 form scales its coefficients by one power of two per matvec:
    scaling the input by 2^k (k = -40, 40) scales the output exactly, and a
    zero input gives a zero output."""
    rng = np.random.default_rng(9)
    p = rng.random((20000, 3))
    q = rng.standard_normal(p.shape)
    f = nnk.KIFMM(nnk.OcTree(p, p, "nodeCapacity", 64), [1e-3, 1.0], "precision", "mixedh")
    u = f.computeVel(q)
    for k in (-40, 40):
        np.testing.assert_array_equal(f.computeVel(q * 2.0 ** k), u * 2.0 ** k)
    assert not f.computeVel(np.zeros_like(q)).any()


def test_mixed_nearest_sphere_pair(oracle, nnk):
    p = np.concatenate([cube_sphere(12), cube_sphere(12, centre=(2.5, 0.3, 0.1))])
    fine = np.concatenate([cube_sphere(48), cube_sphere(48, centre=(2.5, 0.3, 0.1))])
    NN = nnk.nearestMatrix(p, fine)
    q = np.random.default_rng(3).standard_normal(p.shape)
    t = nnk.OcTree(p, p, "nodeCapacity", 40, "nearest", 1, "finePoints", fine, "NN", NN)
    um = nnk.KIFMM(t, [0.01, 1.0], "precision", "mixed").computeVel(q)
    ud = oracle.direct(fine, NN @ q, p, 0.01)
    assert rel(um, ud) < 1e-5


def test_mixed_device_path_and_stats(nnk):
    import torch
    rng = np.random.default_rng(5)
    p = rng.random((30000, 3))
    t = nnk.OcTree(p, p, "nodeCapacity", 64)
    f = nnk.KIFMM(t, [1e-3, 1.0], "precision", "mixed", "svdTol", 1e-9)
    q = rng.standard_normal(p.shape)
    u = f.computeVel(q)
    ud = f.computeVel(torch.tensor(q, device="cuda")).cpu().numpy()
    np.testing.assert_array_equal(u, ud)
    st = np.zeros(14, np.int64)
    assert nnk.lib().nnk_fmm_stats(f._handle().h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 14) == 0
    assert st[12] == 456 and 0 < st[13] <= 456


@pytest.mark.parametrize("precon", [False, True])
def test_refined_solve_reaches_fp64_accuracy(oracle, nnk, precon):
    """Iterative refinement with mixed-form inner solves must meet an FP64
    tolerance on the FP64 system and agree with the plain FP64 GMRES."""
    from importlib import import_module

    from test_gpu_mobility import _system
    sw, B, BT, A1, A2, _ = _system(nnk, oracle, True)
    mob = import_module("stokes-nnkifmm_amd.mobility")
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    x64 = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-10, 300, precon)
    xr = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-10, 600, precon, innerTol=1e-4)
    info = dict(mob.last_solve)
    assert info["flag"] == 0 and info["relres"] <= 1e-10 and info["outer"] >= 2, info
    r = nnk.mobilityMatrixMulti(xr, sw, B, BT) - b
    assert np.linalg.norm(r) <= 1.5e-10 * np.linalg.norm(b)
    assert rel(xr, x64) < 1e-7


@pytest.mark.parametrize("precision,split,res,tol", [("mixedh", 1, 6, 1e-8), ("mixedh", 2, 6, 1e-8),
                                                   ("mixed", 1, 6, 1e-8), ("mixed3", 2, 6, 1e-8),
                                                   ("mixedh", 1, 9, 3e-7)])
def test_mixedh_m2l_kernels_agree_bitwise(nnk, precision, split, res, tol):
    """The split-precision M2L of every level run by the register-ring
    kernel (depth 2 and, but for the three-part form, 3), by the DMA-staged
    kernel (depth 4) and, for the two-part forms, by the DMA-staged kernel
    with fragment prefetch (depth 7: the LDS reads of a chunk overlap the
    products of the previous one), by the DMA-staged kernel
    with four LDS stages (depth 6; with fragment prefetch, 8) and with
    64 x 48 wave tiles (depth 5),
    both on the levels run as class jobs (depth 4 elsewhere): the same
    products in the same order, so the matvecs are bit-identical for one
    split-K factor (with split-K, depth 5's 192-column tiles group the
    partial sums differently: equal to rounding).  For the FP16 form, operator
    rows that are all zero (unscale 0) and rows scaled by widely different
    powers of two come with the clustered points.  Res 6 (R = 428) and the
    headline form's res 9 / 3e-7 (R = 477) leave different parts of the
    last 128-row block as zero padding."""
    rng = np.random.default_rng(44)
    p = np.concatenate([rng.random((30000, 3)), 1e-3 * rng.random((3000, 3)) + 0.5])
    q = rng.standard_normal(p.shape)
    f = nnk.KIFMM(nnk.OcTree(p, p, "nodeCapacity", 48), [1e-3, 1.0], "precision", precision,
                  "coronaRes", res, "svdTol", tol)
    st = np.zeros(14, np.int64)
    assert nnk.lib().nnk_fmm_stats(f._handle().h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 14) == 0
    rank = int(st[13])
    print(f"[m2l] {precision} res {res} tol {tol}: rank {rank}", flush=True)
    if res == 9:  # the upper waves of the last row block hold 1-32 rows
        assert 64 < rank % 128 <= 96
    depths = (2, 4) if precision == "mixed3" else (2, 3, 4, 7)
    lib, h = nnk.lib(), f._handle().h
    i32 = ctypes.POINTER(ctypes.c_int32)
    s0 = np.zeros(16, np.int32)
    d0 = np.zeros(16, np.int32)
    assert lib.nnk_fmm_launch_config(h, s0.ctypes.data_as(i32), d0.ctypes.data_as(i32), 16) == 0
    levels = np.flatnonzero(s0)
    assert len(levels) >= 2 and set(d0[levels]) <= {2, 3, 4}
    out = {}
    for depth in depths:
        s = np.where(s0 > 0, split, 0).astype(np.int32)
        d = np.where(s0 > 0, depth, 0).astype(np.int32)
        assert lib.nnk_fmm_set_launch_config(h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16) == 0
        chk_s = np.zeros(16, np.int32)
        chk_d = np.zeros(16, np.int32)
        lib.nnk_fmm_launch_config(h, chk_s.ctypes.data_as(i32), chk_d.ctypes.data_as(i32), 16)
        assert (chk_s[levels] == split).all() and (chk_d[levels] == depth).all()
        out[depth] = f.computeVel(q)
    if precision != "mixed3":
        # four LDS stages, without and with fragment prefetch; 64 x 48 wave
        # tiles (class-job levels; others keep depth 4)
        for dk in (6, 8, 5):
            wide = []
            for l in levels:
                s = np.zeros(16, np.int32)
                d = np.zeros(16, np.int32)
                s[l], d[l] = split, dk
                if lib.nnk_fmm_set_launch_config(h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16) == 0:
                    wide.append(int(l))
            print(f"[m2l] {precision} split {split}: depth {dk} at levels {wide}", flush=True)
            assert wide
            out[dk] = f.computeVel(q)
    for depth in depths[1:] + ((6, 8) if 6 in out else ()):
        np.testing.assert_array_equal(out[depth], out[2])
    if 5 in out:
        # 192-column tiles split their (larger) active block lists at other
        # points, so with split-K only the grouping of the partial sums differs
        if split == 1:
            np.testing.assert_array_equal(out[5], out[2])
        else:
            assert rel(out[5], out[2]) < 1e-13


@pytest.mark.parametrize("precision", ["mixed", "mixed3", "mixedh"])
def test_pair_jobs_match_class_jobs(nnk, precision):
    """The mixed forms' per-offset pair jobs (sparse levels of an adaptive
    tree: one offset's V-list pairs per job, each block product promoted
    to FP64 on its own, then summed per target in offset-slot order) form
    the same per-block products as the class and merged jobs and add them
    in the same order: bit-identical at split 1, through every kernel that
    runs them (register-ring depth 2/3, DMA-staged 4, DMA-staged with
    fragment prefetch 7)."""
    rng = np.random.default_rng(46)
    p = blobs(rng)
    q = rng.standard_normal(p.shape)
    f = nnk.KIFMM(nnk.OcTree(p, p, "nodeCapacity", 16), [1e-3, 1.0], "precision", precision)
    lib, h = nnk.lib(), f._handle().h
    i32 = ctypes.POINTER(ctypes.c_int32)
    s0 = np.zeros(16, np.int32)
    d0 = np.zeros(16, np.int32)
    assert lib.nnk_fmm_launch_config(h, s0.ctypes.data_as(i32), d0.ctypes.data_as(i32), 16) == 0
    pl = np.flatnonzero(d0 > 10)
    print(f"[pairs] {precision}: per-offset levels {pl.tolist()} of {np.flatnonzero(s0).tolist()}", flush=True)
    assert len(pl) >= 1 and (s0[pl] == 1).all()
    u = f.computeVel(q)

    def setcfg(split, depth):
        s = np.zeros(16, np.int32)
        d = np.zeros(16, np.int32)
        s[pl], d[pl] = split, depth
        return lib.nnk_fmm_set_launch_config(h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16)

    assert setcfg(2, 14) != 0      # no split-K for the pair jobs
    kds = (2, 4) if precision == "mixed3" else (2, 3, 4, 7)
    if precision == "mixed3":
        assert setcfg(1, 13) != 0  # three parts: no third register stage
        assert setcfg(1, 17) != 0  # nor fragment prefetch (its registers would spill)
    for kd in kds:
        assert setcfg(1, 10 + kd) == 0
        np.testing.assert_array_equal(f.computeVel(q), u)
        assert setcfg(1, kd) == 0  # class or merged jobs on the same levels
        np.testing.assert_array_equal(f.computeVel(q), u)


def test_launch_config_rejects_what_a_form_cannot_run(nnk):
    rng = np.random.default_rng(45)
    p = rng.random((20000, 3))
    t = nnk.OcTree(p, p, "nodeCapacity", 64)
    lib = nnk.lib()
    i32 = ctypes.POINTER(ctypes.c_int32)

    def setcfg(f, level, split, depth):
        s = np.zeros(16, np.int32)
        d = np.zeros(16, np.int32)
        s[level], d[level] = split, depth
        return lib.nnk_fmm_set_launch_config(f._handle().h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16)

    def first_m2l_level(f):
        s = np.zeros(16, np.int32)
        d = np.zeros(16, np.int32)
        lib.nnk_fmm_launch_config(f._handle().h, s.ctypes.data_as(i32), d.ctypes.data_as(i32), 16)
        return int(np.flatnonzero(s)[0])

    f64 = nnk.KIFMM(t, [1e-3, 1.0])
    fb = nnk.KIFMM(t, [1e-3, 1.0], "precision", "mixed3")
    fh = nnk.KIFMM(t, [1e-3, 1.0], "precision", "mixedh")
    u64, ub, uh = f64.computeVel(p), fb.computeVel(p), fh.computeVel(p)
    l64, lb, lh = first_m2l_level(f64), first_m2l_level(fb), first_m2l_level(fh)
    assert setcfg(f64, l64, 1, 3) != 0          # the FP64 form has no ring depth
    assert setcfg(fb, lb, 1, 3) != 0            # three parts: no third register stage
    assert setcfg(fb, lb, 1, 5) != 0            # three parts: no 64 x 48 wave tiles
    assert setcfg(fb, lb, 1, 7) != 0            # three parts: no fragment prefetch
    assert setcfg(fh, lh, 1, 9) != 0
    assert setcfg(fh, lh, 0, 4) == 0            # split 0: level left as it is
    assert setcfg(fh, lh, 65, 3) != 0
    assert setcfg(fh, 0, 1, 3) != 0             # level without M2L
    # failed calls change nothing
    np.testing.assert_array_equal(f64.computeVel(p), u64)
    np.testing.assert_array_equal(fb.computeVel(p), ub)
    np.testing.assert_array_equal(fh.computeVel(p), uh)


@pytest.mark.parametrize("case", ["uniform", "blobs", "distinct"])
def test_fp64t_against_direct_and_mixed3(oracle, nnk, case):
    """The FP64 truncated form (the mixed forms' pseudo-inverses and R x R
    translations with FP64 M2L and FP64 direct sums) at 8 points per edge:
    below 1e-6 against the FP64 direct sum on a uniform tree (1e-5 on the
    adaptive ones, whose deep levels the truncation serves less well), and
    the three-part BF16 form (same operators, 24-bit M2L, FP32-math direct
    sums) within 5e-7 of it."""
    rng = np.random.default_rng(41)
    if case == "uniform":
        p = t = rng.random((30000, 3)); cap, eps = 64, 1e-3
    elif case == "blobs":
        p = t = blobs(rng); cap, eps = 16, 1e-3
    else:
        p = rng.random((8000, 3)); t = np.concatenate([rng.random((3000, 3)), p[:1000]]); cap, eps = 40, 5e-3
    tree = nnk.OcTree(p, t, "nodeCapacity", cap)
    q = rng.standard_normal((int(tree.potentials.sum()), 3))
    print(f"[fp64t] case={case} points={len(tree.points)} nodes={tree.nodeCount}", flush=True)
    ut = nnk.KIFMM(tree, [eps, 1.0], "coronaRes", 8, "precision", "fp64t").computeVel(q)
    u3 = nnk.KIFMM(tree, [eps, 1.0], "coronaRes", 8, "precision", "mixed3").computeVel(q)
    tg = tree.getTargetPoints()
    idx = np.sort(rng.choice(len(tg), min(2000, len(tg)), replace=False))
    ud = oracle.direct(tree.getPotentialPoints(), q, tg[idx], eps)
    et = rel(ut[idx], ud)
    print(f"[fp64t] rel err vs direct {et:.3e}, vs mixed3 {rel(u3, ut):.3e}", flush=True)
    assert et < (1e-6 if case == "uniform" else 1e-5)
    assert rel(u3, ut) < 5e-7
