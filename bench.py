"""Benchmark: KIFMM matvec throughput (Mpts/s) on one MI355X.

Workload (the FMM-matvec configuration): N = 2^20 points i.i.d. uniform in
the unit cube, forces i.i.d. N(0,1) in 3 components (numpy PCG64, fixed
seed), eps = 1e-3, mu = 1, nodeCapacity 64, coronaRes 6 / coronaShells 1.
One step = one full KIFMM matvec (upward + downward pass) with inputs and
outputs resident in HBM.  Tree build and operator precompute are setup.

Prints one JSON line.  Besides the throughput it reports:
  roofline      the M2L translation GEMM (the dominant kernel): algorithmic
                FP64 flops (2 * (3Nq)^2 per V-list pair) over its HIP-event
                time in the timed region, against MI355X's FP64 matrix peak;
  cpu_baseline  the CPU oracle (the reference's algorithm restated in C with
                OpenMP) timed on this host on a bounded sample;
  rel_err       ||u - u_direct|| / ||u_direct|| on 4096 sampled targets,
                u_direct from the GPU direct sum (blockcomputation).
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
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--res", type=int, default=6)
    ap.add_argument("--seed", type=int, default=20240611)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-n", type=int, default=1 << 14, help="CPU baseline sample size")
    return ap.parse_args()


def cpu_baseline(args, eps, mu):
    """Time the CPU oracle's matvec on a bounded sample of the workload: the
    same distribution and parameters at N = cpu_n (levels 0-3, ~32 points
    per leaf as at 2^20)."""
    import oracle
    oracle.build()
    rng = np.random.Generator(np.random.PCG64(args.seed + 1))
    p = rng.random((args.cpu_n, 3))
    q = rng.standard_normal((args.cpu_n, 3))
    tree = oracle.Tree(p, p, nodeCapacity=64)
    t0 = time.perf_counter()
    tree.matvec(q, eps, mu, args.res, 1)
    dt = time.perf_counter() - t0
    return {"value": args.cpu_n / dt / 1e6, "unit": "Mpts/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": f"CPU oracle (reference algorithm: per-node surfaces, dense kernel blocks, "
                      f"per-node pivoted LU, OpenMP over the nodes of a level) on N={args.cpu_n} "
                      f"uniform points, same eps/nodeCapacity/coronaRes; {dt:.2f} s for one matvec"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        sys.exit(2)
    import torch
    nnk = importlib.import_module("stokes-nnkifmm_amd")
    lib = nnk.lib()
    eps, mu, N = 1e-3, 1.0, args.n
    rng = np.random.Generator(np.random.PCG64(args.seed))
    pts = rng.random((N, 3))
    f = rng.standard_normal((N, 3))
    t0 = time.perf_counter()
    tree = nnk.OcTree(pts, pts, "nodeCapacity", 64)
    t_tree = time.perf_counter() - t0
    fmm = nnk.KIFMM(tree, [eps, mu], "coronaRes", args.res)
    dev = torch.device("cuda:0")
    fd = torch.tensor(f, device=dev)
    t0 = time.perf_counter()
    h = fmm._handle()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    for _ in range(args.warmup):
        u = fmm.computeVel(fd)
    torch.cuda.synchronize()
    check = lib.nnk_fmm_set_profiling(h.h, 1)
    assert check == 0
    ms = np.zeros(8)
    cnt = np.zeros(8, np.int64)
    lib.nnk_fmm_phase_times(h.h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 8)  # reset
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        u = fmm.computeVel(fd)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    lib.nnk_fmm_phase_times(h.h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 8)
    lib.nnk_fmm_set_profiling(h.h, 0)
    st = np.zeros(13, np.int64)
    lib.nnk_fmm_stats(h.h, st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), 13)
    K = int(st[12])
    ms_step = dt / args.steps * 1e3
    # ---- roofline of the M2L translation GEMM ----
    m2l_flops = 2.0 * K * K * st[0]              # per matvec
    m2l_ms = ms[3] / args.steps                  # per matvec (4 launches)
    launches = cnt[3] / args.steps
    achieved = m2l_flops / (m2l_ms * 1e-3) / 1e12
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "m2l_fp64_hbm_traffic.json")
    if os.path.exists(tfile):
        with open(tfile) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP64_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP64_MFMA_PEAK_TFLOPS, 4), "traffic": traffic,
            "kernel": "k_gemm_gather<160,64,24,1> (M2L)", "launches_per_step": float(launches),
            "flops_per_step": m2l_flops}
    # ---- accuracy on 4096 sampled targets vs the GPU direct sum ----
    idx = np.sort(rng.choice(N, 4096, replace=False))
    tg = torch.tensor(pts[idx], device=dev)
    sd = torch.tensor(pts, device=dev)
    ud = torch.empty((4096, 3), dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = lib.nnk_direct_apply_device(ctypes.c_void_p(sd.data_ptr()), N, ctypes.c_void_p(fd.data_ptr()),
                                     ctypes.c_void_p(tg.data_ptr()), 4096, eps, mu,
                                     ctypes.c_void_p(ud.data_ptr()), stream)
    assert rc == 0
    torch.cuda.synchronize()
    uo = u.cpu().numpy()[idx]
    udn = ud.cpu().numpy()
    rel_err = float(np.linalg.norm(uo - udn) / np.linalg.norm(udn))
    phases = ["strengths", "S2M", "up_gemm", "M2L", "L2L", "P2L", "down_solve", "leaf"]
    out = {
        "metric": "FMM matvec Mpts/s + mobility solve time at 1e-6 rel err, FP32/MFMA vs FP64",
        "value": round(N / (ms_step * 1e-3) / 1e6, 4),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"KIFMM matvec, N={N} uniform points in the unit cube, forces N(0,1), "
                               f"eps={eps}, mu={mu}, nodeCapacity 64, coronaRes {args.res}/shells 1, "
                               f"plain FP64 form",
                   "N": N, "eps": eps, "coronaRes": args.res, "nodeCapacity": 64, "precision": "fp64",
                   "nodes": int(tree.nodeCount), "parallelism": "single GPU"},
        "rel_err_vs_direct": rel_err,
        "phase_ms_per_step": {p: round(float(ms[i] / args.steps), 3) for i, p in enumerate(phases)},
        "setup_s": {"tree": round(t_tree, 3), "operators": round(t_setup, 3)},
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args, eps, mu)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
