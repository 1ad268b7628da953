"""Resistance problems on the KIFMM operator -- host mirror of the
reference's callers of gmres(@(x) fmm.computeVel(x), ...):
Resistance/Rigid_Resistance.m, Resistance/Gen_Resistance.m and the flow
field pipeline of Data Generation/FlowField.m.

The solve runs on the GPU through the mobility handle with no bodies
(nnk_mobility_create with nb6 = 0): GMRES on S f = u with S the KIFMM
matvec (optionally right-preconditioned by the leaf-block solve).
Rigid_Resistance follows the reference's definitions (velocityModes for the
rigid velocity field; F the sum of the (nearest-mapped) forces; M the sum
of x_j x f_j with x_j the NN-weighted quadrature positions, or the points)
but calls the operator it is given (the reference's version calls an
undefined `this` at Rigid_Resistance.m:23).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import c64, check, lib, ptr

last_solve = {}


class _Resistance:
    def __init__(self, fmm, precon):
        self.fmm_h = fmm._handle()
        self.fmm4_h = fmm._handle(res=4) if precon else None
        n = int(fmm.tree.potentials.sum())
        self.N = n
        self._arrs = [np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0),
                      np.zeros(3 * n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0), np.zeros(0), np.zeros(0)]
        a = self._arrs
        h = ctypes.c_void_p()
        check(lib().nnk_mobility_create(self.fmm_h.h, self.fmm4_h.h if self.fmm4_h else None, 0,
                                        ptr(a[0], _lib._i64), ptr(a[1], _lib._i32), ptr(a[2]),
                                        ptr(a[3], _lib._i64), ptr(a[4], _lib._i32), ptr(a[5]),
                                        ptr(a[6]), ptr(a[7]), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().nnk_mobility_destroy(self.h)
            except Exception:
                pass
            self.h = None


def _handle(fmm, precon):
    key = ("_resistance", bool(precon))
    cache = fmm.__dict__.setdefault("_res_cache", {})
    if key not in cache:
        cache[key] = _Resistance(fmm, precon)
    return cache[key]


def resistanceSolve(fmm, vel, tol=1e-6, maxIter=100, precon=False):
    """f = gmres(@(x) fmm.computeVel(x), vel, [], tol, maxIter): vel and f in
    the KIFMM's layout ('GMRES'/'format', KIFMM.m:83-103; (N,3) rows when
    GMRES = 0).  Records flag, relres, iter, resvec in last_solve."""
    r = _handle(fmm, precon)
    rows = fmm._to_rows(c64(vel))
    b = np.ascontiguousarray(rows.T).ravel()          # all x, all y, all z
    x = np.empty_like(b)
    flag = np.zeros(1, np.int32)
    it = np.zeros(1, np.int32)
    rel = np.zeros(1)
    resvec = np.zeros(int(maxIter) + 1)
    check(lib().nnk_mobility_solve(r.h, ptr(b), float(tol), int(maxIter), int(bool(precon)), ptr(x),
                                   ptr(flag, _lib._i32), ptr(rel), ptr(it, _lib._i32), ptr(resvec)))
    last_solve.clear()
    last_solve.update(flag=int(flag[0]), relres=float(rel[0]), iter=int(it[0]),
                      resvec=resvec[:int(it[0]) + 1].copy())
    return fmm._from_rows(x.reshape(3, -1).T, vel)


def velocityModes(points, vel, mom):
    """Rigid velocity U + Omega x x at the points (Geometries/velocityModes.m)."""
    p = np.asarray(points, float).reshape(-1, 3)
    return np.asarray(vel, float).ravel()[None, :] + np.cross(np.asarray(mom, float).ravel()[None, :], p)


def Rigid_Resistance(fmm, vel, mom, tol=1e-6, maxIter=100, precon=False):
    """Total force and moment on the rigid body of fmm's tree moving with
    (vel, mom) (Resistance/Rigid_Resistance.m:1-51): returns F, M, flag,
    relres, iter."""
    tree = fmm.tree
    pts = tree.getPotentialPoints()
    u = velocityModes(pts, vel, mom)
    f = fmm._to_rows(resistanceSolve(fmm, fmm._from_rows(u, np.zeros(3 * len(pts))), tol, maxIter, precon))
    if tree.arguments["nearest"]:
        NN = tree.NN
        F = (NN @ f).sum(axis=0)
        x = NN.T @ tree.getFinePoints()
    else:
        F = f.sum(axis=0)
        x = pts
    M = np.cross(x, f).sum(axis=0)
    return F, M, last_solve["flag"], last_solve["relres"], last_solve["iter"]


def Gen_Resistance(fmm, tol=1e-6, maxIter=100, precon=False):
    """Grand resistance matrix R = [FU FO; MU MO] of the rigid body of fmm's
    tree (Resistance/Gen_Resistance.m:12-26)."""
    cols = []
    e = np.eye(3)
    for k in range(3):
        cols.append(Rigid_Resistance(fmm, e[k], np.zeros(3), tol, maxIter, precon)[:2])
    for k in range(3):
        cols.append(Rigid_Resistance(fmm, np.zeros(3), e[k], tol, maxIter, precon)[:2])
    return np.block([[np.stack([c[0] for c in cols[:3]], 1), np.stack([c[0] for c in cols[3:]], 1)],
                     [np.stack([c[1] for c in cols[:3]], 1), np.stack([c[1] for c in cols[3:]], 1)]])
