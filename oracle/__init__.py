"""CPU oracle for the KIFMM hot path -- TEST INFRASTRUCTURE ONLY.

The oracle restates the MATLAB reference's algorithm (see the header of
nnk_oracle.c for the file:line map) in FP64 C with OpenMP, plus the mobility
layer in numpy (oracle/mobility.py).  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg may import it, and only as the checker or the
timed CPU baseline -- never as part of the product path.

Parity pins (what checks each oracle function independently of itself):
  kernel            closed form: symmetry, r=0 value I/(4 pi mu eps), Oseen
                    far field (tests/test_oracle_kernel.py)
  surfaces          closed-form counts/coordinates (tests/test_oracle_kernel.py)
  tree              hand-worked trees, invariants, root+8 for the sphere case
                    (tests/test_oracle_tree.py)
  lists             complete uniform trees: |U|=27 and |V|=189 in the interior,
                    per-level V totals; U/V symmetry, X = W^T and exactly-once
                    coverage of every (target leaf, source leaf) pair on
                    adaptive trees; W and X lists (contents and order: the
                    re-sorted rest + last merged node of the W coarsening,
                    X's column-major find) of two adaptive trees worked out
                    by hand (tests/hand_trees.py)
  nearestMatrix     brute force over scipy cKDTree, row sums = 1
  fmm matvec        exact agreement with the direct sum when every leaf is
                    adjacent to every other (root + 8 leaves); error vs direct
                    sum decreasing with coronaRes
  reduced pass/solve block-diagonal direct sum; residual of the solve
  mobility/gmres    6 pi mu a U drag on a sphere; numpy.linalg.solve;
                    precon and saddle-point operator against dense matrices
                    written from genMatrixComponets/ReducedSolve/MobilityPrecon
                    on the one-sphere tree (tests/test_oracle_mobility.py)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "_build", "libnnk_oracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i64 = ctypes.POINTER(ctypes.c_int64)
_i32 = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build():
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        L.or_last_error.restype = ctypes.c_char_p
        L.or_kernel.argtypes = [_d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _d]
        L.or_direct.argtypes = [_d, ctypes.c_int64, _d, _d, ctypes.c_int64, ctypes.c_double, ctypes.c_double, _d]
        L.or_surface.argtypes = [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.or_surface_count.argtypes = [ctypes.c_int, ctypes.c_int]
        L.or_lu_solve.argtypes = [_d, ctypes.c_int, _d]
        L.or_tree_build.argtypes = [_d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_void_p)]
        L.or_tree_build_cube.argtypes = [_d, ctypes.c_int64, _d, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _d,
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.or_tree_free.argtypes = [ctypes.c_void_p]
        for f in ("or_tree_npts", "or_tree_npot", "or_tree_ntgt"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int64
        for f in ("or_tree_count", "or_tree_cap"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int32
        L.or_tree_export.argtypes = [ctypes.c_void_p, _d, _u8, _u8, _i32, _d, _i32, _i32, _i32]
        L.or_tree_list_total.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.or_tree_list_total.restype = ctypes.c_int32
        L.or_tree_list_export.argtypes = [ctypes.c_void_p, ctypes.c_int, _i32, _i32]
        L.or_tree_set_nearest.argtypes = [ctypes.c_void_p, _d, ctypes.c_int64, _i64, _i64, _d]
        L.or_nearest_matrix.argtypes = [_d, ctypes.c_int64, _d, ctypes.c_int64, _i64, _i64, _d, ctypes.c_int64]
        L.or_nearest_matrix.restype = ctypes.c_int64
        L.or_fmm_matvec.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_int, _d, _d]
        L.or_reduced_apply.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _d, _d]
        L.or_reduced_solve.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, _d, _d]
        L.or_num_threads.restype = ctypes.c_int
        L.or_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n):
    lib().or_set_threads(int(n))


def num_threads():
    return lib().or_num_threads()


def kernel(x, x0, eps, mu=1.0):
    """Dense (3M, 3N) regularized Stokeslet matrix, rows = targets x0."""
    x = _c(x).reshape(-1, 3)
    x0 = _c(x0).reshape(-1, 3)
    S = np.empty((3 * len(x0), 3 * len(x)))
    lib().or_kernel(_p(x), len(x), _p(x0), len(x0), eps, mu, _p(S))
    return S


def direct(src, q, tgt, eps, mu=1.0):
    src = _c(src).reshape(-1, 3)
    q = _c(q).reshape(-1, 3)
    tgt = _c(tgt).reshape(-1, 3)
    u = np.empty_like(tgt)
    lib().or_direct(_p(src), len(src), _p(q), _p(tgt), len(tgt), eps, mu, _p(u))
    return u


def surface(corners, res, shells=1, down=False):
    corners = _c(corners)
    n = lib().or_surface_count(res, shells)
    out = np.empty((n, 3))
    lib().or_surface(_p(corners), res, shells, int(bool(down)), _p(out))
    return out


def lu_solve(A, b):
    A = _c(A).copy()
    b = _c(b).copy()
    lib().or_lu_solve(_p(A), A.shape[0], _p(b))
    return b


def nearest_matrix(coarse, fine):
    """nearestMatrix: returns (rows, cols, weights) 0-based COO, column-major."""
    coarse = _c(coarse).reshape(-1, 3)
    fine = _c(fine).reshape(-1, 3)
    cap = max(4 * len(fine), 16)
    while True:
        r = np.empty(cap, np.int64)
        c = np.empty(cap, np.int64)
        w = np.empty(cap)
        k = lib().or_nearest_matrix(_p(coarse), len(coarse), _p(fine), len(fine),
                                    _p(r, _i64), _p(c, _i64), _p(w), cap)
        if k >= 0:
            return r[:k], c[:k], w[:k]
        cap = -k


class Tree:
    """Oracle octree (OcTree.m semantics); ids are 1-based like MATLAB."""

    def __init__(self, pot, tgt, nodeCapacity=200, maxDepth=21, finePoints=None, NN=None, rootCube=None):
        pot = _c(pot).reshape(-1, 3)
        tgt = _c(tgt).reshape(-1, 3)
        h = ctypes.c_void_p()
        cube = None if rootCube is None else _c(rootCube).ravel()
        rc = lib().or_tree_build_cube(_p(pot), len(pot), _p(tgt), len(tgt), int(nodeCapacity), int(maxDepth),
                                      None if cube is None else _p(cube), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        self._h = h
        L = lib()
        self.npts = L.or_tree_npts(h)
        self.npot = L.or_tree_npot(h)
        self.ntgt = L.or_tree_ntgt(h)
        self.nodeCount = L.or_tree_count(h)
        self.nodeCapacity = L.or_tree_cap(h)
        n = self.nodeCount
        self.points = np.empty((self.npts, 3))
        self.potentials = np.empty(self.npts, np.uint8)
        self.targets = np.empty(self.npts, np.uint8)
        self.pointIndex = np.empty(self.npts, np.int32)
        self.nodeCorners = np.empty((n, 6))
        self.nodeLevel = np.empty(n, np.int32)
        self.nodeParents = np.empty(n, np.int32)
        self.nodeChildren = np.empty((n, 8), np.int32)
        L.or_tree_export(h, _p(self.points), _p(self.potentials, _u8), _p(self.targets, _u8),
                         _p(self.pointIndex, _i32), _p(self.nodeCorners), _p(self.nodeLevel, _i32),
                         _p(self.nodeParents, _i32), _p(self.nodeChildren, _i32))
        self.potentials = self.potentials.astype(bool)
        self.targets = self.targets.astype(bool)
        self.lists = []
        for q in range(4):
            tot = L.or_tree_list_total(h, q)
            ptr = np.empty(n + 1, np.int32)
            idx = np.empty(max(tot, 1), np.int32)
            L.or_tree_list_export(h, q, _p(ptr, _i32), _p(idx, _i32))
            self.lists.append((ptr, idx[:tot]))
        self.nearest = False
        if finePoints is not None:
            self.set_nearest(finePoints, NN)

    def set_nearest(self, fine, NN):
        """NN: scipy.sparse (Q x npot) or None to build it with nearestMatrix."""
        import scipy.sparse as sp
        fine = _c(fine).reshape(-1, 3)
        if NN is None:
            r, c, w = nearest_matrix(self.points[self.potentials], fine)
            NN = sp.csc_matrix((w, (r, c)), shape=(len(fine), self.npot))
        NN = sp.csc_matrix(NN)
        NN.sort_indices()
        colptr = _c(NN.indptr, np.int64)
        row = _c(NN.indices, np.int64)
        val = _c(NN.data)
        lib().or_tree_set_nearest(self._h, _p(fine), len(fine), _p(colptr, _i64), _p(row, _i64), _p(val))
        self.finePoints = fine
        self.NN = NN
        self.nearest = True

    def interactions(self, node, kind):
        """kind 0..3 = U, V, W, X; node is 1-based."""
        ptr, idx = self.lists[kind]
        return idx[ptr[node - 1]:ptr[node]]

    def matvec(self, q, eps, mu=1.0, res=6, shells=1):
        q = _c(q).reshape(-1, 3)
        assert len(q) == self.npot
        u = np.zeros((self.ntgt, 3))
        lib().or_fmm_matvec(self._h, eps, mu, res, shells, _p(q), _p(u))
        return u

    def reduced_apply(self, q, eps, mu=1.0):
        q = _c(q).reshape(-1, 3)
        u = np.zeros((self.ntgt, 3))
        lib().or_reduced_apply(self._h, eps, mu, _p(q), _p(u))
        return u

    def reduced_solve(self, v, eps, mu=1.0):
        v = _c(v).reshape(-1, 3)
        f = np.zeros((self.npot, 3))
        rc = lib().or_reduced_solve(self._h, eps, mu, _p(v), _p(f))
        if rc != 0:
            raise RuntimeError(lib().or_last_error().decode())
        return f

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib().or_tree_free(self._h)
                self._h = None
        except Exception:
            pass
