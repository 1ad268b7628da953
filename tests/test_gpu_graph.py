"""HIP-graph replay of the matvec and of the GMRES operator (GPU).

The graph replays the same kernels in the same order as the eager path, so
results must be bit-identical: the matvec is replayed several times with
different inputs and eager work in between, each replay checked against the
eager result; a mobility solve with the graph must reproduce the eager
solve exactly (same iterates, same answer)."""
import numpy as np
import pytest

from geometry import blobs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixedh"])
def test_matvec_graph_replays_match_eager(nnk, precision):
    import torch
    rng = np.random.default_rng(11)
    p = blobs(rng)
    t = nnk.OcTree(p, p, "nodeCapacity", 24)
    eager = nnk.KIFMM(t, [2e-3, 1.0], "precision", precision)
    graph = nnk.KIFMM(t, [2e-3, 1.0], "precision", precision, "hipGraph", 1)
    for it in range(4):
        q = torch.tensor(rng.standard_normal(p.shape), device="cuda")
        ug = graph.computeVel(q)
        ue = eager.computeVel(q)
        torch.cuda.synchronize()
        assert torch.equal(ug, ue), it
    # host path through the same graph
    qh = rng.standard_normal(p.shape)
    np.testing.assert_array_equal(graph.computeVel(qh), eager.computeVel(qh))


@pytest.mark.parametrize("precon", [False, True])
def test_mobility_solve_graph_matches_eager(oracle, nnk, precon):
    from importlib import import_module

    from test_gpu_mobility import _system
    mob = import_module("stokes-nnkifmm_amd.mobility")
    out = []
    for graph in (0, 1):
        sw, B, BT, A1, A2, _ = _system(nnk, oracle, True)
        sw.FMM.arguments["hipGraph"] = graph
        b = mob._rhs(sw, lambda t, x: np.zeros_like(x), 0.0)
        x = mob.mobility_gmres(sw, B, BT, A1, A2, b, 1e-10, 300, precon)
        out.append((x, dict(mob.last_solve)))
    (x0, i0), (x1, i1) = out
    assert i0["iter"] == i1["iter"] and i0["flag"] == i1["flag"] == 0
    np.testing.assert_array_equal(x0, x1)


@pytest.mark.parametrize("precision", ["fp64", "mixed3", "mixedh"])
def test_autotuned_launch_shapes(nnk, precision):
    """The autotuner only reorders sums: tuned results agree with the
    default launch shapes to rounding, are deterministic once tuned, and
    the chosen per-level shapes are reported."""
    import ctypes
    rng = np.random.default_rng(12)
    p = rng.random((60000, 3))
    t = nnk.OcTree(p, p, "nodeCapacity", 32)
    q = rng.standard_normal(p.shape)
    u0 = nnk.KIFMM(t, [1e-3, 1.0], "precision", precision).computeVel(q)
    f = nnk.KIFMM(t, [1e-3, 1.0], "precision", precision, "autotune", 1)
    u1 = f.computeVel(q)
    u2 = f.computeVel(q)
    np.testing.assert_array_equal(u1, u2)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.linalg.norm(u1 - u0) <= tol * np.linalg.norm(u0)
    split = np.zeros(16, np.int32)
    depth = np.zeros(16, np.int32)
    i32 = ctypes.POINTER(ctypes.c_int32)
    assert nnk.lib().nnk_fmm_launch_config(f._handle().h, split.ctypes.data_as(i32), depth.ctypes.data_as(i32),
                                           16) == 0
    assert (split[2:] >= 0).all() and split[2] >= 1
    # FP64: 0 class tiles, 1 per-offset pair jobs; mixed: ring depth 2/3,
    # DMA-staged 4-8, per-offset pair jobs 12-14, 17
    assert set(depth[split > 0]) <= ({0, 1} if precision == "fp64" else {2, 3, 4, 5, 6, 7, 8, 12, 13, 14, 17})
    threads = ctypes.c_int32()
    assert nnk.lib().nnk_fmm_leaf_threads(f._handle().h, ctypes.byref(threads)) == 0
    assert threads.value in (128, 256, 512)
