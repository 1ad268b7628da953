"""Mobility problem -- host mirror of the reference's Mobility/*.m.

genMatrixComponets assembles B, BT, A1, A2 exactly as the reference does
(scipy.sparse standing in for MATLAB sparse).  mobilityMatrixMulti,
mobilityPrecon and the GMRES solve run on the GPU through nnk_mobility_*;
the vector layout is the reference's: [f (3N: all x, all y, all z);
U (3 nsw); Omega (3 nsw)].
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import c64, check, lib, ptr


def genMatrixComponets(swimmers):
    """[B, BT, A1, A2] (Mobility/genMatrixComponets.m:16-100)."""
    import scipy.sparse as sp
    nearest = bool(swimmers.Tree.arguments["nearest"])
    _, _, NN, points, fine = swimmers.getSwimmers()
    Nsw = points.shape[0]
    NBnd = swimmers.getBnd()[0].shape[0]
    N = Nsw + NBnd
    nsw = swimmers.swimmerNo()
    sizes, _ = swimmers.getSwimmerSizes()
    si = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    rows = np.arange(Nsw)
    body = np.repeat(np.arange(nsw), sizes)
    au = sp.csc_matrix((-np.ones(Nsw), (rows, body)), shape=(N, nsw))
    AU = sp.kron(sp.identity(3), au, format="csc")
    # rotated (not translated) coordinates of each swimmer's points
    rot = np.concatenate([swimmers.getIndSwimmer(n + 1)[0] for n in range(nsw)]) if nsw else np.zeros((0, 3))
    xm = [sp.csc_matrix((rot[:, d], (rows, body)), shape=(N, nsw)) for d in range(3)]
    Ze = sp.csc_matrix((N, nsw))
    AOm = sp.bmat([[Ze, -xm[2], xm[1]], [xm[2], Ze, -xm[0]], [-xm[1], xm[0], Ze]], format="csc")
    if nearest:
        NN = sp.csc_matrix(NN)
        w = np.asarray(NN.sum(axis=0)).ravel()[:Nsw]          # column sums of NN
        pm = (NN.T @ fine)[:Nsw]                               # NN-weighted fine positions
    else:
        w = np.ones(Nsw)
        pm = points
    af = sp.csr_matrix((w, (body, rows)), shape=(nsw, N))
    AF = sp.kron(sp.identity(3), af, format="csr")
    xm2 = [sp.csr_matrix((pm[:, d], (body, rows)), shape=(nsw, N)) for d in range(3)]
    Z2 = sp.csr_matrix((nsw, N))
    AM = sp.bmat([[Z2, -xm2[2], xm2[1]], [xm2[2], Z2, -xm2[0]], [-xm2[1], xm2[0], Z2]], format="csr")
    B = sp.vstack([AF, AM], format="csr")
    BT = sp.hstack([AU, AOm], format="csr")
    A1 = (B @ B.T).tocsr()
    A2 = (BT.T @ BT).tocsr()
    return B, BT, A1, A2


class _MobHandle:
    def __init__(self, swimmers, B, BT, A1, A2, precon):
        import scipy.sparse as sp
        fmm = swimmers.FMM
        self.fmm_h = fmm._handle()
        self.fmm4_h = fmm._handle(res=4) if precon else None
        B, BT = sp.csr_matrix(B), sp.csr_matrix(BT)
        B.sort_indices()
        BT.sort_indices()
        self.keep = (B, BT)
        nb6 = B.shape[0]
        self.n = B.shape[1] + nb6
        A1 = np.ascontiguousarray(A1.toarray() if sp.issparse(A1) else A1, dtype=np.float64)
        A2 = np.ascontiguousarray(A2.toarray() if sp.issparse(A2) else A2, dtype=np.float64)
        arrs = [np.ascontiguousarray(B.indptr, np.int64), np.ascontiguousarray(B.indices, np.int32), c64(B.data),
                np.ascontiguousarray(BT.indptr, np.int64), np.ascontiguousarray(BT.indices, np.int32),
                c64(BT.data), A1, A2]
        self._arrs = arrs
        h = ctypes.c_void_p()
        check(lib().nnk_mobility_create(self.fmm_h.h, self.fmm4_h.h if self.fmm4_h else None, nb6,
                                        ptr(arrs[0], _lib._i64), ptr(arrs[1], _lib._i32), ptr(arrs[2]),
                                        ptr(arrs[3], _lib._i64), ptr(arrs[4], _lib._i32), ptr(arrs[5]),
                                        ptr(arrs[6]), ptr(arrs[7]), ctypes.byref(h)))
        self.h = h
        self.ids = (id(B), id(BT))
        if fmm.arguments.get("hipGraph"):
            check(lib().nnk_mobility_set_graph(h, 1))

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().nnk_mobility_destroy(self.h)
            except Exception:
                pass
            self.h = None


def _mobility(swimmers, B, BT, A1=None, A2=None, precon=False):
    m = getattr(swimmers, "_mobility", None)
    if m is not None and m.src == (id(B), id(BT)) and (m.fmm4_h is not None or not precon) \
            and m.fmm is swimmers.FMM:
        return m
    if A1 is None:
        A1 = B @ B.T
    if A2 is None:
        A2 = BT.T @ BT
    m = _MobHandle(swimmers, B, BT, A1, A2, precon)
    m.src = (id(B), id(BT))
    m.fmm = swimmers.FMM
    swimmers._mobility = m
    return m


def mobilityMatrixMulti(f, swimmers, B, BT):
    """[S f + BT [U;Om]; B f] (Mobility/MobilityMatrixMulti.m:17-36)."""
    m = _mobility(swimmers, B, BT)
    x = c64(f).ravel()
    y = np.empty_like(x)
    check(lib().nnk_mobility_apply(m.h, ptr(x), ptr(y)))
    return y


def mobilityPrecon(f, swimmers, B, BT, A1, A2, tol=None, maxIter=None):
    """Block preconditioner (Mobility/MobilityPrecon.m:17-39); tol and
    maxIter are unused, as in the reference."""
    m = _mobility(swimmers, B, BT, A1, A2, precon=True)
    x = c64(f).ravel()
    y = np.empty_like(x)
    check(lib().nnk_mobility_precon(m.h, ptr(x), ptr(y)))
    return y


last_solve = {}


def mobility_gmres(swimmers, B, BT, A1, A2, rhs, tol, maxIter, precon):
    """x = gmres(afunc, rhs, [], tol, maxIter) with afunc = A or A*P and
    x = P(x) afterwards (mobilityProblem.m:47-58).  Also records flag,
    relres, iter and resvec in mobility.last_solve."""
    m = _mobility(swimmers, B, BT, A1, A2, precon=bool(precon))
    b = c64(rhs).ravel()
    x = np.empty_like(b)
    flag = np.zeros(1, np.int32)
    it = np.zeros(1, np.int32)
    rel = np.zeros(1)
    resvec = np.zeros(int(maxIter) + 1)
    check(lib().nnk_mobility_solve(m.h, ptr(b), float(tol), int(maxIter), int(bool(precon)), ptr(x),
                                   ptr(flag, _lib._i32), ptr(rel), ptr(it, _lib._i32), ptr(resvec)))
    last_solve.clear()
    last_solve.update(flag=int(flag[0]), relres=float(rel[0]), iter=int(it[0]),
                      resvec=resvec[:int(it[0]) + 1].copy())
    return x


def mobility_refined(swimmers, B, BT, A1, A2, rhs, tol, maxIter, precon, innerTol=1e-4, svdTol=1e-8,
                     innerPrecision="mixed", outer="refine", innerRes=None):
    """The solve of mobility_gmres around a mixed-precision inner solver (no
    counterpart in the reference): GMRES on the mixed-precision KIFMM of the
    same tree to innerTol (preconditioned with mixed-precision coronaRes-4
    operators too) inside an FP64 outer iteration, either iterative
    refinement (outer="refine": FP64 residual, inner correction, add) or
    flexible GMRES (outer="fgmres": FP64 Arnoldi on the FP64 operator, each
    basis vector preconditioned by an inner solve).  innerRes sets the
    corona resolution of the inner operator (default: the FP64 operator's);
    a coarser one makes every inner iteration cheaper, and the outer
    iteration, on the FP64 operator, still decides convergence.  The answer
    meets tol on the unpreconditioned FP64 system.  Records flag, relres,
    iter (inner iterations) and outer in mobility.last_solve."""
    if outer not in ("refine", "fgmres"):
        raise ValueError(f"mobility_refined: unknown outer iteration {outer!r}")
    from .kifmm import _Handle
    m = _mobility(swimmers, B, BT, A1, A2, precon=bool(precon))
    fmm = swimmers.FMM
    a = fmm.arguments
    res = int(innerRes) if innerRes is not None else int(a["coronaRes"])
    key = (innerPrecision, float(svdTol), bool(precon), res)
    if not hasattr(m, "inner"):
        m.inner = {}
    inner = m.inner.get(key)
    if inner is None:
        tune = bool(a.get("autotune"))
        inner = (_Handle(fmm.tree, a["kernelPar"][0], a["kernelPar"][1], res, a["coronaShells"],
                         innerPrecision, svdTol, autotune=tune),
                 _Handle(fmm.tree, a["kernelPar"][0], a["kernelPar"][1], 4, a["coronaShells"],
                         innerPrecision, svdTol, autotune=tune) if precon else None)
        m.inner[key] = inner
    b = c64(rhs).ravel()
    x = np.empty_like(b)
    flag = np.zeros(1, np.int32)
    it = np.zeros(1, np.int32)
    rel = np.zeros(1)
    nout = np.zeros(1, np.int32)
    solve = lib().nnk_mobility_solve_fgmres if outer == "fgmres" else lib().nnk_mobility_solve_refined
    check(solve(m.h, inner[0].h, inner[1].h if inner[1] else None, ptr(b), float(tol), int(maxIter),
                float(innerTol), int(bool(precon)), ptr(x), ptr(flag, _lib._i32), ptr(rel), ptr(it, _lib._i32),
                ptr(nout, _lib._i32)))
    last_solve.clear()
    last_solve.update(flag=int(flag[0]), relres=float(rel[0]), iter=int(it[0]), outer=int(nout[0]))
    return x


def _rhs(swimmers, fluidFunc, t):
    pointVel = swimmers.getVelocities()
    pts = swimmers.getSwimmers()[0]
    fluidVel = np.asarray(fluidFunc(t, pts), float).reshape(-1, 3)
    vel = pointVel.copy()
    vel[:len(fluidVel)] -= fluidVel
    force, moment = swimmers.getForceMoment()
    return np.concatenate([vel.T.ravel(), force[0::3], force[1::3], force[2::3], moment[0::3],
                           moment[1::3], moment[2::3]])


def _unpack(swimmers, x, nsw):
    out = np.zeros(9 * nsw)
    U = x[len(x) - 6 * nsw:len(x) - 3 * nsw].reshape(3, nsw).T
    Om = x[len(x) - 3 * nsw:].reshape(3, nsw).T
    for i in range(nsw):
        b1, b2 = swimmers.b[i]
        out[9 * i:9 * i + 3] = U[i]
        out[9 * i + 3:9 * i + 6] = np.cross(Om[i], b1)
        out[9 * i + 6:9 * i + 9] = np.cross(Om[i], b2)
    return out


def mobilityProblem(swimmers, fluidFunc, t, conditions, kernelPar, TreePar, threads, GPU, blockSize,
                    tol, maxIter, PreconSettings):
    """Velocities and basis rates [U; Om x b1; Om x b2] per body
    (Mobility/mobilityProblem.m:24-72)."""
    conditions = np.asarray(conditions, float).ravel()
    nsw = len(conditions) // 9
    for i in range(nsw):
        d = conditions[9 * i:9 * i + 9]
        swimmers.updateSwimmmer(i + 1, d[0:3], d[3:9])
    return _solve(swimmers, fluidFunc, t, kernelPar, TreePar, threads, GPU, blockSize, tol, maxIter,
                  PreconSettings, nsw)


def mobilityProblemGravity(swimmers, fluidFunc, t, conditions, kernelPar, TreePar, threads, GPU,
                           blockSize, tol, maxIter, PreconSettings, gravPar):
    """As mobilityProblem with F = G and M = -(4/3) pi offset (b3 x G)
    (Mobility/mobilityProblemGravity.m:25-74)."""
    conditions = np.asarray(conditions, float).ravel()
    gravPar = np.asarray(gravPar, float).ravel()
    nsw = len(conditions) // 9
    for i in range(nsw):
        d = conditions[9 * i:9 * i + 9]
        swimmers.updateSwimmmer(i + 1, d[0:3], d[3:9])
        swimmers.updateSwimmmerRHS(i + 1, gravPar[0:3], gravPar[3])
    return _solve(swimmers, fluidFunc, t, kernelPar, TreePar, threads, GPU, blockSize, tol, maxIter,
                  PreconSettings, nsw)


def _solve(swimmers, fluidFunc, t, kernelPar, TreePar, threads, GPU, blockSize, tol, maxIter,
           PreconSettings, nsw):
    swimmers.genTree(*TreePar)
    swimmers.genKIFMM(kernelPar, "parThreads", threads, "GPU", GPU, "GMRES", 1, "format", 2,
                      "blockSize", blockSize)
    B, BT, A1, A2 = genMatrixComponets(swimmers)
    rhs = _rhs(swimmers, fluidFunc, t)
    precon = PreconSettings is not None and len(PreconSettings) > 0
    x = mobility_gmres(swimmers, B, BT, A1, A2, rhs, tol, maxIter, precon)
    return _unpack(swimmers, x, nsw)
