"""nearestMatrix -- host mirror of Tools/Nearest/nearestMatrix.m.

NN(q, j) = 1/#ties for the coarse point(s) j nearest to fine point q,
computed on the GPU by nnk_nearest_matrix; returned as a scipy CSC matrix
(Q x N) like MATLAB's sparse().  blockSize and threads are accepted for
signature compatibility.
"""
import numpy as np

from . import _lib
from ._lib import c64, check, lib, ptr


def nearestMatrix(coarse, fine, blockSize=0.2, threads=0):
    import scipy.sparse as sp
    coarse = c64(coarse).reshape(-1, 3)
    fine = c64(fine).reshape(-1, 3)
    n, q = len(coarse), len(fine)
    cap = max(2 * q, 16)
    while True:
        colptr = np.empty(n + 1, np.int64)
        row = np.empty(cap, np.int64)
        val = np.empty(cap)
        nnz = np.zeros(1, np.int64)
        rc = lib().nnk_nearest_matrix(ptr(coarse), n, ptr(fine), q, ptr(colptr, _lib._i64),
                                      ptr(row, _lib._i64), ptr(val), cap, ptr(nnz, _lib._i64))
        if rc == 0:
            k = int(nnz[0])
            return sp.csc_matrix((val[:k], row[:k], colptr), shape=(q, n))
        if nnz[0] > cap:
            cap = int(nnz[0])
            continue
        check(rc)
