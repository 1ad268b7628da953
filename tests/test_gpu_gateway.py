"""The MATLAB boundary end to end (GPU): a C program built with gcc drives
libnnkifmm through the MEX gateway's marshalling layer
(stokes-nnkifmm_amd/matlab/nnk_gateway.c) with the data exactly as MATLAB
hands it over -- (3,n) column-major points, MATLAB CSC with mwIndex rows,
B and BT as the CSC of B.' and BT.', A1.' / A2.' -- on the one-sphere
mobility problem (384 force / 6,144 quadrature points, eps 0.01).  Its
tree export, matvecs, preconditioner and GMRES solution must equal the
Python host path's bit for bit; the program also checks the handle
registry's lifetimes (deferred release, refusal of released ids,
mexAtExit clean-up)."""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from geometry import cube_sphere

pytestmark = pytest.mark.gpu
IDX = np.uint64   # mwIndex (size_t)


def _w(d, name, a, dtype):
    np.ascontiguousarray(a, dtype=dtype).tofile(os.path.join(d, name))


def _r(d, name):
    return np.fromfile(os.path.join(d, name), dtype=np.float64)


def test_gateway_matches_python_path(nnk, tmp_path):
    exe = os.path.join(os.path.dirname(nnk.library_path()), "nnk_gateway_test")
    assert os.path.exists(exe), "build() makes the gateway test program"
    eps, mu, tol, cap = 0.01, 1.0, 1e-6, 64
    c, fine = cube_sphere(8), cube_sphere(32)
    NN = nnk.nearestMatrix(c, fine)
    sw = nnk.Swimmers()
    sw.addSwimmer(c, fine, NN, [0, 0, 0], [1, 0, 0, 0, 1, 0], np.zeros_like(c), [6 * np.pi, 0, 0], [0, 0, 0])
    sw.genTree("nodeCapacity", cap, "nearest", 1)
    sw.genKIFMM([eps, mu], "GMRES", 1, "format", 2)
    B, BT, A1, A2 = nnk.genMatrixComponets(sw)
    pts, fpts, NNall = sw.getSwimmerBnd()
    n, q, nb6 = len(pts), len(fpts), B.shape[0]
    rng = np.random.default_rng(12)
    f = rng.standard_normal((n, 3))
    x = rng.standard_normal(3 * n + nb6)
    NNc = sp.csc_matrix(NNall)
    NNc.sort_indices()
    Br, BTr = sp.csr_matrix(B), sp.csr_matrix(BT)     # CSR of B = CSC of B.'
    Br.sort_indices()
    BTr.sort_indices()
    d = str(tmp_path)
    _w(d, "sizes.i64", [n, q, NNc.nnz, nb6, Br.nnz, BTr.nnz, cap], np.int64)
    _w(d, "pot3.f64", pts, np.float64)                # (3,n) column-major = n x 3 row-major
    _w(d, "fine3.f64", fpts, np.float64)
    _w(d, "nn_jc.idx", NNc.indptr, IDX)
    _w(d, "nn_ir.idx", NNc.indices, IDX)
    _w(d, "nn_pr.f64", NNc.data, np.float64)
    _w(d, "f3.f64", f, np.float64)
    _w(d, "bt_jc.idx", Br.indptr, IDX)
    _w(d, "bt_ir.idx", Br.indices, IDX)
    _w(d, "bt_pr.f64", Br.data, np.float64)
    _w(d, "btt_jc.idx", BTr.indptr, IDX)
    _w(d, "btt_ir.idx", BTr.indices, IDX)
    _w(d, "btt_pr.f64", BTr.data, np.float64)
    _w(d, "a1t.f64", np.asarray(A1.todense()).T.ravel(order="F"), np.float64)   # A1.' column-major
    _w(d, "a2t.f64", np.asarray(A2.todense()).T.ravel(order="F"), np.float64)
    _w(d, "x.f64", x, np.float64)
    _w(d, "params.f64", [eps, mu, tol], np.float64)
    print("running", exe, flush=True)
    r = subprocess.run([exe, d], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0 and "GATEWAY_OK" in r.stdout, r.stdout

    # the Python host path on the same inputs
    t = sw.Tree
    nc = t.nodeCount
    np.testing.assert_array_equal(_r(d, "points.f64"), t.points.ravel(order="F"))
    np.testing.assert_array_equal(_r(d, "corners.f64"), t.nodeCorners.ravel(order="F"))
    np.testing.assert_array_equal(_r(d, "children.f64"), t.nodeChildren.ravel(order="F"))
    np.testing.assert_array_equal(_r(d, "level.f64"), t.nodeLevel)
    np.testing.assert_array_equal(_r(d, "pointindex.f64"), t.pointIndex)
    for kind in range(4):
        ptr, idx = _r(d, f"list{kind}_ptr.f64"), _r(d, f"list{kind}_idx.f64")
        assert len(ptr) == nc + 1
        for node in range(1, nc + 1):
            np.testing.assert_array_equal(idx[int(ptr[node - 1]):int(ptr[node])], t.interactions(node, kind + 1))
    fmm = nnk.KIFMM(t, [eps, mu])          # (n,3) in and out
    np.testing.assert_array_equal(_r(d, "u_apply.f64"), fmm.computeVel(f).ravel())
    np.testing.assert_array_equal(_r(d, "u_reduced.f64"), fmm.reducedComputeVel(f).ravel())
    np.testing.assert_array_equal(_r(d, "y_apply.f64"), nnk.mobilityMatrixMulti(x, sw, B, BT))
    np.testing.assert_array_equal(_r(d, "y_precon.f64"), nnk.mobilityPrecon(x, sw, B, BT, A1, A2, tol, 100))
    mob = nnk.mobility
    xs = mob.mobility_gmres(sw, B, BT, A1, A2, x, tol, 200, True)
    info = mob.last_solve
    flag, relres, iters = _r(d, "solve.f64")
    assert (flag, iters) == (info["flag"], info["iter"]) and relres == info["relres"]
    np.testing.assert_array_equal(_r(d, "x_solve.f64"), xs)
