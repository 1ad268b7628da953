"""Benchmark: KIFMM matvec throughput (Mpts/s) on one MI355X.

Workload (the FMM-matvec configuration): N = 2^20 points i.i.d. uniform in
the unit cube, forces i.i.d. N(0,1) in 3 components (numpy PCG64, fixed
seed), eps = 1e-3, mu = 1, nodeCapacity 64, coronaShells 1.
One step = one full KIFMM matvec (upward + downward pass) with inputs and
outputs resident in HBM.  Tree build and operator precompute are setup.
Forms timed (all reported under "forms"):
  fp64_res6    the plain FP64 form at the reference's coronaRes 6
  fp64_res8    the same at coronaRes 8 (its error does not improve: the
               exact check-to-equivalent inverses are ill-conditioned)
  mixed_res6   truncated-SVD pseudo-inverses, two-part split-BF16 M2L on the
               matrix cores, FP32-math direct sums
  mixed3_res8  the same with three-part split-BF16 M2L at coronaRes 8 (below
               1e-6 relative error)
  mixed3_res9_tol3e-7
               three-part split-BF16 M2L on the headline form's operators
               (coronaRes 9, cut 3e-7, rank 477)
  mixedh_res8  two-part split-FP16 M2L on power-of-two scaled operands at
               coronaRes 8: the same accuracy with half the MFMAs
  fp64t_res8   the truncated-SVD operators of mixed3_res8 (cut 1e-8,
               rank 605) with every operation in FP64: FP64-MFMA R x R M2L
               and FP64 direct sums (the FP64 comparator below 1e-6)
  fp64t_res9_tol3e-7
               the same on the headline form's operators (rank 477)
  mixedh_res9_tol3e-7
               the same at coronaRes 9 with the pseudo-inverses cut at
               3e-7 (rank 477 instead of 605 at 8 / 1e-8, below one
               128-row tile boundary): the fastest form below 1e-6 in a
               sweep of corona resolution and cut, the headline `value`

Prints one JSON line.  Besides the throughput it reports:
  roofline      the headline form's M2L translation GEMM (the dominant
                kernel): its MFMA flops (FP64: 2 (3Nq)^2 per V-list pair;
                mixed: 3 BF16 products of 2 R^2 per pair, R the
                pseudo-inverse rank) over its HIP-event time in the timed
                region, against the MI355X dense MFMA peak of that type;
                roofline.matvec: the whole matvec, sum over the kernel
                classes of (algorithmic flops / that class's peak) over the
                measured time per matvec, with the per-class terms;
  cpu_baseline  the CPU oracle (the reference's algorithm restated in C with
                OpenMP) timed on this host on a bounded sample;
  rel_err       ||u - u_direct|| / ||u_direct|| on 4096 sampled targets,
                u_direct from the GPU direct sum (blockcomputation).
  mobility solves (skipped with --no-solve), each with wall time,
  iterations and residuals: one sphere (384 / 6144 points, FP64 GMRES to
  1e-6); 64 spheroids above a wall (~126k / ~2M points, FP64 and
  mixed-refined); 256 batched single-body problems (FP64 and FP32-math,
  HIP-graph replay); the squirmer-pair adjoint (forward, batched transpose
  GMRES, gradient kernels).
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_MFMA_PEAK_TFLOPS = 78.6
BF16_MFMA_PEAK_TFLOPS = 2500.0
FP64_VALU_PEAK_TFLOPS = 78.6
FP32_VALU_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
PAIR_FLOPS = 33   # one regularized-Stokeslet source-target interaction


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--res", type=int, default=6, help="corona resolution of the CPU baseline")
    ap.add_argument("--svd-tol", type=float, default=1e-8, help="pseudo-inverse cut of the mixed forms")
    ap.add_argument("--forms", default="fp64_res6,fp64_res8,fp64t_res8,fp64t_res9_tol3e-7,mixed_res6,mixed3_res8,"
                                      "mixed3_res9_tol3e-7,mixedh_res8,mixedh_res9_tol3e-7",
                    help="comma-separated forms <precision>_res<r>[_tol<svd tol>] to time; the last one is the "
                         "headline")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the mobility solves (one sphere, spheroids + wall, batched, adjoint)")
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default M2L launch shapes (no timing of candidates in the warm-up)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-n", type=int, default=1 << 16, help="CPU baseline sample size (all host threads)")
    ap.add_argument("--cpu-n1", type=int, default=1 << 12, help="CPU baseline sample size on one thread")
    ap.add_argument("--solve-only", choices=["sphere", "spheroids", "batched", "adjoint", "ode45"], default=None,
                    help="run only this mobility solve and print its record (for profiling runs)")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB (3 times): a kernel of "
                         "known traffic for scaling the FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def cpu_baseline(args, eps, mu):
    """Time the CPU oracle's matvec on a bounded sample of the workload: the
    same distribution and parameters at N = cpu_n on every host thread, and
    at N = cpu_n1 on one thread."""
    import oracle
    oracle.build()
    allc = oracle.num_threads()

    def one(n, threads):
        oracle.set_threads(threads)
        rng = np.random.Generator(np.random.PCG64(args.seed + 1))
        p = rng.random((n, 3))
        q = rng.standard_normal((n, 3))
        tree = oracle.Tree(p, p, nodeCapacity=64)
        t0 = time.perf_counter()
        tree.matvec(q, eps, mu, args.res, 1)
        return time.perf_counter() - t0
    dt = one(args.cpu_n, allc)
    dt1 = one(args.cpu_n1, 1)
    oracle.set_threads(allc)
    what = ("CPU oracle (reference algorithm: per-node surfaces, dense kernel blocks, per-node pivoted LU, "
            "OpenMP over the nodes of a level), same distribution, eps, nodeCapacity 64, coronaRes "
            f"{args.res}")
    return {"value": args.cpu_n / dt / 1e6, "unit": "Mpts/s", "cores": allc, "kind": "port",
            "sample": f"{what}; N={args.cpu_n} on {allc} threads: {dt:.2f} s for one matvec",
            "single_core": {"value": args.cpu_n1 / dt1 / 1e6, "unit": "Mpts/s", "cores": 1,
                            "sample": f"N={args.cpu_n1} on 1 thread: {dt1:.2f} s for one matvec"}}


def _oracle_system(sw, eps, cap):
    """The oracle twin (oracle/mobility.py) of a Swimmers system."""
    import scipy.sparse as sp
    import oracle
    from oracle.mobility import MobilityOracle
    pts, fpts, NNall = sw.getSwimmerBnd()
    t = oracle.Tree(pts, pts, nodeCapacity=cap, finePoints=fpts, NN=NNall)
    _, _, NNs, rot, frot = sw.getSwimmers()
    sizes, _ = sw.getSwimmerSizes()
    ptr = np.concatenate([[0], np.cumsum(sizes)])
    w = np.asarray(sp.csc_matrix(NNs).sum(0)).ravel()
    pm = sp.csc_matrix(NNs).T @ frot
    return MobilityOracle(t, eps, 1.0, ptr, rot, w, pm)


def cpu_solves(nnk, gpu):
    """The solves on the CPU oracle (the reference's algorithm, every host
    thread).  The one-sphere solve runs in full.  The larger ones are timed
    per GMRES iteration (one saddle-point apply and one preconditioner
    application) on the stated or a labelled reduced system, and the solve
    time is estimated as that time x the GPU run's iteration count (x the
    point ratio for the reduced suspension)."""
    import importlib as il
    import oracle
    oracle.build()
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    out = {"threads": oracle.num_threads(), "kind": "port"}
    # one sphere (384 / 6144 points, eps 0.01), preconditioned GMRES to 1e-6
    sw = nnk.Swimmers()
    c, fine = geo.cube_sphere(8), geo.cube_sphere(32)
    sw.addSwimmer(c, fine, nnk.nearestMatrix(c, fine), [0, 0, 0], [1, 0, 0, 0, 1, 0], np.zeros_like(c),
                  [6 * np.pi, 0, 0], [0, 0, 0])
    m = _oracle_system(sw, 0.01, 64)
    b = np.zeros(m.n)
    b[m.n3:m.n3 + 3] = [6 * np.pi, 0, 0]
    t0 = time.perf_counter()
    x, info = m.gmres(b, 1e-6, 200, True)
    dt = time.perf_counter() - t0
    g = gpu.get("mobility_solve_one_sphere", {}).get("precon", {})
    out["one_sphere"] = {"wall_s": round(dt, 3), "iterations": int(info["iter"]), "U_x": float(x[m.n3]),
                         "gpu_wall_s": g.get("wall_s"), "gpu_iterations": g.get("iterations")}

    def per_iteration(m):
        x = np.random.default_rng(1).standard_normal(m.n)
        t0 = time.perf_counter()
        m.apply(m.precon(x))
        return time.perf_counter() - t0
    # spheroid suspension, reduced: 8 spheroids (384 / 6144 points) above a
    # 12 x 12 wall in an 8^3 box, eps 0.005
    sw = geo.spheroid_suspension(nbody=8, n=8, wall_n=12, box=8.0)
    m = _oracle_system(sw, 0.005, 64)
    it = per_iteration(m)
    g = gpu.get("mobility_solve_spheroids_wall", {})
    n_full = g.get("force_points")
    gi = g.get("fp64_precon", {}).get("iterations")
    rec = {"system": f"8 spheroids + 12 x 12 wall, {m.N} force points (reduced from the stated 64 + 40 x 40)",
           "s_per_iteration": round(it, 3)}
    if n_full and gi:
        rec["estimated_full_solve_s"] = round(it * n_full / m.N * gi, 1)
        rec["estimate"] = (f"s per iteration x {n_full}/{m.N} points x {gi} GPU FP64 iterations "
                           "(the oracle's cost grows about linearly with the points at fixed leaf size)")
        rec["gpu_fp64_wall_s"] = g.get("fp64_precon", {}).get("wall_s")
    out["spheroids_wall_reduced"] = rec
    # squirmer pair (2 x 1536 / 24576 points, eps 0.01, B2 = -5) at its stated size
    sw = geo.squirmer_pair(n=16, refine=4, gap=0.5, B1=1.0, B2=-5.0)
    m = _oracle_system(sw, 0.01, 64)
    it = per_iteration(m)
    a = gpu.get("adjoint_squirmer_pair", {})
    gi = max(a.get("adjoint_iterations") or [0])
    out["squirmer_pair"] = {"force_points": int(m.N), "s_per_iteration": round(it, 3),
                            "estimated_forward_solve_s": round(it * gi, 1) if gi else None,
                            "estimated_central_difference_s": round(2 * it * gi, 1) if gi else None,
                            "estimate": "s per iteration x the GPU solve's iteration count; one central "
                                        "difference is two forward solves, and the gradient with respect "
                                        f"to all {3 * m.N} node coordinates by differences takes "
                                        f"{6 * m.N} of them (the adjoint: one transpose solve)",
                            "gpu_forward_s": a.get("forward_s"), "gpu_transpose_gmres_s": a.get("transpose_gmres_s")}
    # batched single bodies: the problem of median size, per iteration
    bodies, eps = geo.batch_bodies()
    sizes = np.array([len(b.Geometries[0]) for b in bodies])
    k = int(np.argsort(sizes)[len(sizes) // 2])
    m = _oracle_system(bodies[k], float(eps[k]), 64)
    it = per_iteration(m)
    bt = gpu.get("batched_solves", {})
    npts = np.array([len(b.Geometries[0]) for b in bodies], float)
    mi = bt.get("fp64", {}).get("mean_iterations")
    rec = {"median_problem_points": int(m.N), "s_per_iteration": round(it, 3)}
    if mi:
        rec["estimated_batch_s"] = round(it * float((npts / m.N).sum()) * mi, 1)
        rec["estimate"] = ("s per iteration of the median problem x sum over the 256 problems of "
                           "their point ratio x the GPU's mean iteration count")
        rec["gpu_fp64_wall_s"] = bt.get("fp64", {}).get("wall_s")
    out["batched_256"] = rec
    return out


def run_form(nnk, lib, tree, fd, args, eps, mu, precision, res, svd_tol):
    """Time one form of the matvec; returns (u, record)."""
    import torch
    P = ctypes.POINTER
    fmm = nnk.KIFMM(tree, [eps, mu], "coronaRes", res, "precision", precision, "svdTol", svd_tol,
                    "autotune", 0 if args.no_autotune else 1)
    t0 = time.perf_counter()
    h = fmm._handle()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    for _ in range(args.warmup):
        u = fmm.computeVel(fd)
    torch.cuda.synchronize()
    assert lib.nnk_fmm_set_profiling(h.h, 1) == 0
    ms = np.zeros(8)
    cnt = np.zeros(8, np.int64)
    lib.nnk_fmm_phase_times(h.h, ms.ctypes.data_as(P(ctypes.c_double)), cnt.ctypes.data_as(P(ctypes.c_int64)), 8)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        u = fmm.computeVel(fd)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lib.nnk_fmm_phase_times(h.h, ms.ctypes.data_as(P(ctypes.c_double)), cnt.ctypes.data_as(P(ctypes.c_int64)), 8)
    lib.nnk_fmm_set_profiling(h.h, 0)
    st = np.zeros(14, np.int64)
    lib.nnk_fmm_stats(h.h, st.ctypes.data_as(P(ctypes.c_int64)), 14)
    spl = np.zeros(16, np.int32)
    dep = np.zeros(16, np.int32)
    lib.nnk_fmm_launch_config(h.h, spl.ctypes.data_as(P(ctypes.c_int32)), dep.ctypes.data_as(P(ctypes.c_int32)), 16)
    leaf_threads = ctypes.c_int32()
    lib.nnk_fmm_leaf_threads(h.h, ctypes.byref(leaf_threads))
    K, R = int(st[12]), int(st[13])
    ms_step = dt / args.steps * 1e3
    phases = ["strengths", "S2M", "up", "M2L", "L2L", "P2L", "down", "leaf"]
    if precision in ("fp64", "fp64t"):
        r = K if precision == "fp64" else R
        flops = 2.0 * r * r * st[0]
        pairs = any(int(dep[l]) == 1 for l in range(16) if spl[l] > 0)
        peak, kern = FP64_MFMA_PEAK_TFLOPS, "k_gemm_gather<160,64,16,1> (M2L, FP64 MFMA" + \
            ("; per-offset pair jobs + k_pair_sum at some levels)" if pairs else ")")
    else:
        np_, f16, what = {"mixed": (2, "false", "two-part split-BF16"), "mixed3": (3, "false", "three-part split-BF16"),
                          "mixedh": (2, "true", "two-part split-FP16")}[precision]
        flops = (6 if np_ == 3 else 3) * 2.0 * R * R * st[0]   # 16-bit products per split product
        peak = BF16_MFMA_PEAK_TFLOPS
        used = {int(dep[l]) for l in range(16) if spl[l] > 0}
        names = ([f"k_m2l_split<{np_},D,1,{f16}>"] if used & {2, 3, 12, 13} else []) + \
                ([f"k_m2l_dma<{np_},1,{f16}>"] if used & {4, 14} else []) + \
                ([f"k_m2l_dma<{np_},1,{f16},3> (64 x 48 wave tiles)"] if 5 in used else []) + \
                ([f"k_m2l_dma<{np_},1,{f16},2,4> (four LDS stages)"] if 6 in used else []) + \
                ([f"k_m2l_dma<{np_},1,{f16},2,3,true> (fragment prefetch)"] if used & {7, 17} else []) + \
                ([f"k_m2l_dma<{np_},1,{f16},2,4,true> (fragment prefetch, four LDS stages)"] if 8 in used else []) + \
                (["per-offset pair jobs + k_pair_sum at some levels"] if used & {12, 13, 14, 17} else [])
        kern = " / ".join(names) + f" (M2L, {what} MFMA{', LDS-DMA staged' if used <= {4, 5, 6, 7, 8} else ''})"
    m2l_ms = ms[3] / args.steps
    achieved = flops / (m2l_ms * 1e-3) / 1e12 if m2l_ms > 0 else 0.0
    rec = {"precision": precision, "coronaRes": res, "svd_tol": svd_tol if precision != "fp64" else None,
           "rank": R if precision != "fp64" else K,
           "matvec_roofline": matvec_roofline(precision, st, K, R, ms_step), "Mpts_s": round(tree.points.shape[0] / (ms_step * 1e-3) / 1e6, 4),
           "ms_per_step": round(ms_step, 3), "setup_s": round(t_setup, 3),
           "phase_ms_per_step": {p: round(float(ms[i] / args.steps), 3) for i, p in enumerate(phases)},
           "m2l_launch_per_level": {str(l): {"split": int(spl[l]), "depth": int(dep[l])}
                                    for l in range(16) if spl[l] > 0},
           "leaf_kernel_threads": int(leaf_threads.value),
           "m2l": {"kernel": kern, "flops_per_step": flops, "ms_per_step": round(m2l_ms, 3),
                   "launches_per_step": float(cnt[3] / args.steps), "achieved_tflops": round(achieved, 3),
                   "peak_tflops": peak}}
    return u, rec


def matvec_roofline(precision, st, K, R, ms_step):
    """Sum over the matvec's kernel classes of (algorithmic flops / that
    class's peak), over the measured time per matvec.  Work counts from
    nnk_fmm_stats: 0 M2L pairs, 1 P2P pairs, 2 S2M pairs, 3 L2T pairs,
    4 M2T pairs, 5 P2L pairs, 6 up-solve nodes, 7 down-solve nodes, 8 M2M
    parents, 9 L2L nodes.  FP64 form: every translation and solve is a
    (3Nq)^2 FP64 GEMM per node (two factors per solve).  Truncated forms:
    R x R translations and the leaves' R x 3Nq up expansions (the down and
    W expansions are not counted); M2L on 16-bit parts counts every split
    product (3 or 6); direct sums at 33 flop per pair on the VALU (FP32
    math in the mixed forms)."""
    f64m, b16 = FP64_MFMA_PEAK_TFLOPS, BF16_MFMA_PEAK_TFLOPS
    v64, v32 = FP64_VALU_PEAK_TFLOPS, FP32_VALU_PEAK_TFLOPS
    direct_peak = v32 if precision in ("mixed", "mixed3", "mixedh") else v64
    pairs = float(st[1] + st[2] + st[3] + st[4] + st[5])
    cls = {"direct_sums (P2P, S2M, L2T, M2T, P2L)": (PAIR_FLOPS * pairs, direct_peak)}
    if precision == "fp64":
        k2 = 2.0 * K * K
        cls["M2L"] = (k2 * st[0], f64m)
        cls["M2M + L2L"] = (k2 * (8 * st[8] + st[9]), f64m)
        cls["equivalent-density solves"] = (2 * k2 * (st[6] + st[7]), f64m)
    else:
        nprod = {"mixed": 3, "mixed3": 6, "mixedh": 3, "fp64t": 1}[precision]
        cls["M2L"] = (nprod * 2.0 * R * R * st[0], f64m if precision == "fp64t" else b16)
        cls["M2M + L2L"] = (2.0 * R * R * (8 * st[8] + st[9]), f64m)
        # up expansion R x 3Nq per leaf with sources (nodes with an up
        # density minus M2M parents); the down and W expansions (3Nq x R
        # per target leaf / W node) are left out, which can only lower frac
        cls["up expansion (leaves)"] = (2.0 * R * K * (st[6] - st[8]), f64m)
    ideal_ms = sum(f / (pk * 1e12) for f, pk in cls.values()) * 1e3
    return {"frac": round(ideal_ms / ms_step, 4), "ideal_ms": round(ideal_ms, 3),
            "classes": {k: {"flops": f, "peak_tflops": pk, "ideal_ms": round(f / (pk * 1e12) * 1e3, 4)}
                        for k, (f, pk) in cls.items()}}


def handle_ideal_ms(lib, h, precision):
    """Ideal time (ms) of one matvec of a KIFMM handle: the roofline sum of
    matvec_roofline over its work counts."""
    P = ctypes.POINTER
    st = np.zeros(14, np.int64)
    lib.nnk_fmm_stats(h.h, st.ctypes.data_as(P(ctypes.c_int64)), 14)
    return matvec_roofline(precision, st, int(st[12]), int(st[13]), 1.0)["ideal_ms"]


def direct_sample(lib, pts, fd, idx, eps, mu):
    import torch
    dev = fd.device
    tg = torch.tensor(pts[idx], device=dev)
    sd = torch.tensor(pts, device=dev)
    ud = torch.empty((len(idx), 3), dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.nnk_direct_apply_device(ctypes.c_void_p(sd.data_ptr()), len(pts), ctypes.c_void_p(fd.data_ptr()),
                                       ctypes.c_void_p(tg.data_ptr()), len(idx), eps, mu,
                                       ctypes.c_void_p(ud.data_ptr()), stream) == 0
    torch.cuda.synchronize()
    return ud.cpu().numpy()


def sphere_solve(nnk):
    """One rigid sphere (radius 1, mu 1, eps 0.01), 384 force / 6144 NEAREST
    quadrature points, driven by F = 6 pi e_x: GMRES to 1e-6 (FP64)."""
    import importlib as il
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    mob = il.import_module("stokes-nnkifmm_amd.mobility")
    c, fine = geo.cube_sphere(8), geo.cube_sphere(32)
    out = {}
    for pre in (False, True):
        sw = nnk.Swimmers()
        t0 = time.perf_counter()
        NN = nnk.nearestMatrix(c, fine)
        sw.addSwimmer(c, fine, NN, [0, 0, 0], [1, 0, 0, 0, 1, 0], np.zeros_like(c), [6 * np.pi, 0, 0], [0, 0, 0])
        U = nnk.mobilityProblem(sw, lambda t, x: np.zeros_like(x), 0.0, np.r_[0, 0, 0, 1, 0, 0, 0, 1, 0],
                                [0.01, 1.0], ["nodeCapacity", 64, "nearest", 1], 0, 1, 0.1, 1e-6, 200,
                                [1e-6, 100] if pre else [])
        dt = time.perf_counter() - t0
        info = mob.last_solve
        out["precon" if pre else "plain"] = {"wall_s": round(dt, 4), "iterations": info["iter"],
                                             "relres": info["relres"], "U_x": float(U[0]),
                                             "stokes_law_rel_err": abs(float(U[0]) - 1.0)}
    return out


def spheroid_solve(nnk):
    """64 prolate spheroids (aspect 3, semi-major axis 1) above a 40 x 40
    wall in a 20^3 box, 4x refined NEAREST quadrature (~126k force / ~2M
    quadrature points), eps = 0.005, each body pulled by a unit force:
    GMRES to 1e-6 with the block preconditioner, FP64, by iterative
    refinement around the mixed form (FP64 residuals) and by flexible GMRES
    (FP64 Arnoldi, mixed-form inner solves to 1e-2)."""
    import torch
    import importlib as il
    nvtx = torch.cuda.nvtx   # ROCTx ranges: a kernel trace splits by them
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    mob = il.import_module("stokes-nnkifmm_amd.mobility")
    nvtx.range_push("spheroids: set-up")
    t0 = time.perf_counter()
    sw = geo.spheroid_suspension()
    t_geo = time.perf_counter() - t0
    t0 = time.perf_counter()
    sw.genTree("nodeCapacity", 64, "nearest", 1)
    sw.genKIFMM([0.005, 1.0], "GMRES", 1, "format", 2, "autotune", 1)
    B, BT, A1, A2 = nnk.genMatrixComponets(sw)
    t_setup = time.perf_counter() - t0
    b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
    out = {"force_points": int(sw.Tree.points.shape[0]), "quadrature_points": int(len(sw.getSwimmerBnd()[1])),
           "tree_nodes": int(sw.Tree.nodeCount), "setup_s": {"geometry_nearest": round(t_geo, 3),
                                                             "tree_B": round(t_setup, 3)}}
    # nearestMatrix over the whole suspension (every coarse point against
    # every fine point, 2.6e11 pairs by brute force); the grid search is exact
    pts, fine_all, _ = sw.getSwimmerBnd()
    t0 = time.perf_counter()
    NNg = nnk.nearestMatrix(pts, fine_all)
    t_nn = time.perf_counter() - t0
    out["nearest_matrix_whole_system"] = {
        "coarse": int(len(pts)), "fine": int(len(fine_all)), "wall_s": round(t_nn, 3), "nnz": int(NNg.nnz),
        "pairs_brute_force": float(len(pts)) * float(len(fine_all))}
    torch.cuda.synchronize()
    nvtx.range_pop()
    nvtx.range_push("spheroids: FP64 solve")
    t0 = time.perf_counter()
    x64 = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-6, 400, True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nvtx.range_pop()
    info = mob.last_solve
    # roofline of the solve: per iteration one coronaRes-6 and one
    # coronaRes-4 matvec at their ideal (roofline-sum) times; the leaf-block
    # solves, SpMVs and Arnoldi vectors are not counted (a lower bound)
    lib = nnk.lib()
    i6 = handle_ideal_ms(lib, sw.FMM._handle(), "fp64")
    i4 = handle_ideal_ms(lib, sw.FMM._handle(res=4), "fp64")
    ideal_s = (info["iter"] + 2) * (i6 + i4) * 1e-3
    out["fp64_precon"] = {"wall_s": round(dt, 3), "iterations": info["iter"], "relres": info["relres"],
                          "roofline": {"frac": round(ideal_s / dt, 4), "ideal_s": round(ideal_s, 4),
                                       "matvec_ideal_ms": {"coronaRes6": i6, "coronaRes4": i4},
                                       "counted": "iterations + 2 applications of the res-6 and res-4 "
                                                  "FP64 matvecs (roofline sum over kernel classes); "
                                                  "leaf-block solves, SpMVs and Arnoldi not counted"}}
    nvtx.range_push("spheroids: refinement solve")
    t0 = time.perf_counter()
    xr = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-6, 400, True, innerTol=1e-3)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nvtx.range_pop()
    info = mob.last_solve
    r = nnk.mobilityMatrixMulti(xr, sw, B, BT) - b
    nb = 6 * sw.swimmerNo()
    out["mixed_refined_precon"] = {
        "wall_s": round(dt, 3), "inner_iterations": info["iter"], "outer": info["outer"],
        "relres_fp64": float(np.linalg.norm(r) / np.linalg.norm(b)),
        "body_velocity_diff_vs_fp64": float(np.linalg.norm(xr[-nb:] - x64[-nb:]) / np.linalg.norm(x64[-nb:]))}
    nvtx.range_push("spheroids: flexible GMRES solve")
    t0 = time.perf_counter()
    xf = mob.mobility_refined(sw, B, BT, A1, A2, b, 1e-6, 400, True, innerTol=1e-2, outer="fgmres")
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nvtx.range_pop()
    info = mob.last_solve
    r = nnk.mobilityMatrixMulti(xf, sw, B, BT) - b
    # roofline of the solve: each inner iteration one mixed coronaRes-6 and
    # one mixed coronaRes-4 matvec, each outer iteration the FP64 pair, at
    # their ideal (roofline-sum) times (a lower bound, as for the FP64 solve)
    inner6, inner4 = sw._mobility.inner[("mixed", 1e-8, True, 6)]
    m6, m4 = handle_ideal_ms(lib, inner6, "mixed"), handle_ideal_ms(lib, inner4, "mixed")
    ideal_f = (info["iter"] * (m6 + m4) + (info["outer"] + 2) * (i6 + i4)) * 1e-3
    out["mixed_fgmres_precon"] = {
        "wall_s": round(dt, 3), "inner_iterations": info["iter"], "outer": info["outer"], "inner_tol": 1e-2,
        "roofline": {"frac": round(ideal_f / dt, 4), "ideal_s": round(ideal_f, 4),
                     "matvec_ideal_ms": {"mixed_coronaRes6": m6, "mixed_coronaRes4": m4},
                     "counted": "inner iterations x the mixed res-6 and res-4 matvecs + (outer + 2) x the "
                                "FP64 pair (roofline sum over kernel classes); leaf-block solves, SpMVs and "
                                "Arnoldi not counted"},
        "relres_fp64": float(np.linalg.norm(r) / np.linalg.norm(b)),
        "body_velocity_diff_vs_fp64": float(np.linalg.norm(xf[-nb:] - x64[-nb:]) / np.linalg.norm(x64[-nb:]))}
    return out


def batched_solves(nnk):
    """256 independent single-body problems (spheres and spheroids, 200-5000
    force points log-uniform, 4x refined quadrature, eps in [0.002, 0.02]):
    one batched GMRES to 1e-6, HIP-graph replay of the persistent batched
    operator with its P2P workgroup size autotuned, FP64 and FP32-math."""
    import importlib as il
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    t0 = time.perf_counter()
    bodies, eps = geo.batch_bodies()
    out = {"problems": len(bodies), "force_points": int(sum(len(b.Geometries[0]) for b in bodies)),
           "setup_geometry_s": round(time.perf_counter() - t0, 3)}
    x64 = None
    for prec in ("fp64", "mixed"):
        t0 = time.perf_counter()
        bm = nnk.BatchedMobility(bodies, eps, precision=prec, hipGraph=True, autotune=True)
        b = bm.rhs()
        bm.apply(b)   # the autotuner's timing runs and the graph capture belong to set-up
        ts = time.perf_counter() - t0
        t0 = time.perf_counter()
        x = bm.solve(b, 1e-6, 300)
        dt = time.perf_counter() - t0
        i = bm.last_solve
        rec = {"setup_s": round(ts, 3), "wall_s": round(dt, 3), "p2p_waves_autotuned": bm.p2pWaves(),
               "max_iterations": int(i["iter"].max()),
               "mean_iterations": float(i["iter"].mean()), "unconverged": int(i["flag"].sum()),
               "max_relres": float(i["relres"].max())}
        if x64 is None:
            x64 = x
        else:
            rec["rel_diff_vs_fp64"] = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
        out[prec] = rec
    return out


def adjoint_pair(nnk):
    """Squirmer pair (radius 1, gap 0.5, B1 = 1, B2 = -5; 1536 force and
    24576 quadrature points each, eps = 0.01): forward solve, one batched
    transpose GMRES for the six translational-velocity functionals and the
    adjoint gradients with respect to every node position and eps."""
    import importlib as il
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    t0 = time.perf_counter()
    sw = geo.squirmer_pair(n=16, refine=4, gap=0.5, B1=1.0, B2=-5.0)
    bm6 = nnk.BatchedMobility([sw] * 6, [0.01] * 6)
    ts = time.perf_counter() - t0
    n = bm6.n // 6
    N = int(bm6.npts[0])
    t0 = time.perf_counter()
    z = bm6.solve(bm6.rhs(), 1e-10, 400)
    t_fwd = time.perf_counter() - t0
    U = z[3 * N:3 * N + 6].reshape(3, 2).T
    e = np.zeros(6 * n)
    for k in range(6):   # U of body k % 2, component k // 2
        e[k * n + 3 * N + (k // 2) * 2 + k % 2] = 1.0
    t0 = time.perf_counter()
    lam = bm6.solve(e, 1e-10, 400, transpose=True)
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    gx, gy, ge = bm6.gradient(z, lam)
    t_grad = time.perf_counter() - t0
    return {"setup_s": round(ts, 3), "forward_s": round(t_fwd, 3), "transpose_gmres_s": round(t_adj, 3),
            "gradient_kernels_s": round(t_grad, 3), "U_body0": U[0].tolist(), "U_body1": U[1].tolist(),
            "dU_deps": ge.tolist(), "max_abs_dU_dx": float(max(np.abs(g).max() for g in gx)),
            "adjoint_iterations": bm6.last_solve["iter"].tolist()}


def time_stepping_squirmers(nnk):
    """Data Generation/Squiremers.m: 10 squirmers (radius 0.5, B1 = 2) in a
    closed box of six 10 x 10 walls (spacing 0.3), eps = 0.1, nodeCapacity
    500, GMRES tol 1e-4 (at most 50 iterations), preconditioned; ode45
    (RelTol 1e-3) over t in [0, 0.3].  Every right-hand side rebuilds the
    tree, the KIFMM handles and B/BT as the reference does; the per-level
    operators come from the operator cache when it is on (the walls fix the
    root box, so every rebuild hits)."""
    import gc
    import importlib as il

    import torch
    geo = il.import_module("stokes-nnkifmm_amd.geometries")
    t0 = time.perf_counter()
    sw, y0 = geo.squirmer_box(nsw=10, seed=1)
    t_geo = time.perf_counter() - t0
    calls = [0]

    def rhs(t, x):
        calls[0] += 1
        return nnk.mobilityProblem(sw, lambda t, p: np.zeros_like(p), t, x, [0.1, 1.0],
                                   ["nodeCapacity", 500, "nearest", 1], 20, 0, 1, 1e-4, 50, [1e-5, 100])

    def timed(k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            r = rhs(0.0, y0)
        return (time.perf_counter() - t0) / k, r
    nnk.operatorCache(0)
    gc.collect()
    cold, r_cold = timed(3)
    nnk.operatorCache(8 << 30)
    rhs(0.0, y0)
    warm, r_warm = timed(3)
    assert np.array_equal(r_cold, r_warm)
    calls[0] = 0
    t0 = time.perf_counter()
    t, y = nnk.ode45(rhs, [0.0, 0.3], y0)
    t_ode = time.perf_counter() - t0
    hits, misses, kept, _ = nnk.operatorCacheStats()
    nnk.operatorCache(0)
    disp = np.linalg.norm((y[-1] - y[0]).reshape(-1, 9)[:, :3], axis=1)
    return {"bodies": 10, "force_points": int(len(sw.getSwimmerBnd()[0])), "setup_geometry_s": round(t_geo, 3),
            "rhs_cold_s": round(cold, 4), "rhs_warm_s": round(warm, 4), "rhs_speedup": round(cold / warm, 2),
            "ode45": {"t_end": 0.3, "wall_s": round(t_ode, 3), "rhs_evaluations": calls[0],
                      "output_points": int(len(t)), "mean_displacement": float(disp.mean())},
            "cache": {"hits": hits, "misses": misses, "bytes_kept": kept},
            "bit_identical_cold_vs_warm": True}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_mb > 0:
        n = args.copy_mb << 17   # doubles in copy_mb MiB
        a = torch.ones(n, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n}), flush=True)
        return
    nnk = importlib.import_module("stokes-nnkifmm_amd")
    lib = nnk.lib()
    if args.solve_only:
        fn = {"sphere": sphere_solve, "spheroids": spheroid_solve, "batched": batched_solves,
              "adjoint": adjoint_pair, "ode45": time_stepping_squirmers}[args.solve_only]
        print(json.dumps({"solve": args.solve_only, "record": fn(nnk)}), flush=True)
        return
    eps, mu, N = 1e-3, 1.0, args.n
    rng = np.random.Generator(np.random.PCG64(args.seed))
    pts = rng.random((N, 3))
    f = rng.standard_normal((N, 3))
    nnk.OcTree(pts[:4096], pts[:4096], "nodeCapacity", 64)   # GPU runtime and code objects loaded
    t0 = time.perf_counter()
    tree = nnk.OcTree(pts, pts, "nodeCapacity", 64)
    t_tree = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    fd = torch.tensor(f, device=dev)
    idx = np.sort(rng.choice(N, 4096, replace=False))
    ud = direct_sample(lib, pts, fd, idx, eps, mu)
    forms = {}
    names = [f for f in args.forms.split(",") if f]
    for name in names:
        parts = name.split("_")
        prec, res = parts[0], int(parts[1][3:])
        tol = float(parts[2][3:]) if len(parts) > 2 else args.svd_tol
        if prec not in ("fp64", "fp64t", "mixed", "mixed3", "mixedh") or not parts[1].startswith("res"):
            raise SystemExit(f"unknown form {name}")
        u, rec = run_form(nnk, lib, tree, fd, args, eps, mu, prec, res, 0.0 if prec == "fp64" else tol)
        uo = u.cpu().numpy()[idx]
        rec["rel_err_vs_direct"] = float(np.linalg.norm(uo - ud) / np.linalg.norm(ud))
        forms[name] = rec
        print(f"# {name}: {rec['Mpts_s']} Mpts/s, {rec['ms_per_step']} ms/step, "
              f"rel err {rec['rel_err_vs_direct']:.3e}", file=sys.stderr, flush=True)
        del u
        torch.cuda.empty_cache()
    head = forms[names[-1]]
    m = head["m2l"]
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"m2l_{names[-1]}_hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    roof = {"bound": "mfma", "achieved": m["achieved_tflops"], "peak": m["peak_tflops"], "unit": "TFLOP/s",
            "frac": round(m["achieved_tflops"] / m["peak_tflops"], 4), "traffic": traffic,
            "kernel": m["kernel"], "launches_per_step": m["launches_per_step"],
            "flops_per_step": m["flops_per_step"], "matvec": head["matvec_roofline"]}
    out = {
        "metric": "FMM matvec Mpts/s + mobility solve time at 1e-6 rel err, FP32/MFMA vs FP64",
        "value": head["Mpts_s"],
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "bf16x2-split-mfma/f32/f64", "mixed3": "bf16x3-split-mfma/f32/f64",
                  "mixedh": "fp16x2-split-mfma/f32/f64"}.get(
            head["precision"], "f64"),
        "data": "synthetic",
        "config": {"workload": f"KIFMM matvec, N={N} uniform points in the unit cube, forces N(0,1), "
                               f"eps={eps}, mu={mu}, nodeCapacity 64",
                   "N": N, "eps": eps, "nodeCapacity": 64, "nodes": int(tree.nodeCount),
                   "headline_form": names[-1], "coronaRes": head["coronaRes"],
                   "parallelism": "single GPU"},
        "rel_err_vs_direct": head["rel_err_vs_direct"],
        "forms": forms,
        "setup_s": {"tree": round(t_tree, 3),
                    "tree_build_ms": dict(zip(["dedup", "subdivide", "point_lists", "uvwx_lists"],
                                              [round(float(v), 2) for v in tree.buildTimes]))},
        "roofline": roof,
    }
    for name in names[:-1]:
        if name.startswith("fp64"):
            out[f"speedup_vs_{name}"] = round(forms[name]["ms_per_step"] / head["ms_per_step"], 3)
    if not args.no_solve:
        out["mobility_solve_one_sphere"] = sphere_solve(nnk)
        print("# sphere solve done", file=sys.stderr, flush=True)
        out["mobility_solve_spheroids_wall"] = spheroid_solve(nnk)
        print("# spheroid solve done", file=sys.stderr, flush=True)
        out["batched_solves"] = batched_solves(nnk)
        print("# batched solves done", file=sys.stderr, flush=True)
        out["adjoint_squirmer_pair"] = adjoint_pair(nnk)
        print("# adjoint done", file=sys.stderr, flush=True)
        out["time_stepping_squirmers"] = time_stepping_squirmers(nnk)
        print("# time stepping done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, eps, mu)
        print("# cpu matvec baseline done", file=sys.stderr, flush=True)
        if not args.no_solve:
            out["cpu_baseline"]["solves"] = cpu_solves(nnk, out)
            print("# cpu solve baselines done", file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
