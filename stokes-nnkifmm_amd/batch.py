"""Batched single-body mobility problems (no counterpart in the reference,
which runs one mobilityProblem per body: Mobility/mobilityProblem.m:24-72).

BatchedMobility takes a list of Swimmers systems (any bodies, optional
fixed boundary, nearest quadrature or none) with their eps, and solves all
their mobility problems at once on the GPU (libnnkifmm nnk_batch_*): one
persistent P2P launch per matvec over (problem, target tile) work items,
lockstep GMRES, optional HIP-graph replay, the transpose operator and
adjoint gradients with respect to node positions and eps.  Per problem the
system is the reference's [S BT; B 0] with S the direct Stokeslet sum (what
the KIFMM reduces to when the tree is the root and 8 leaves).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import c64, check, lib, ptr

_PREC = {"fp64": 0, "mixed": 1}


class BatchedMobility:
    def __init__(self, systems, eps, mu=1.0, precision="fp64", hipGraph=True, autotune=False):
        import scipy.sparse as sp
        if precision not in _PREC:
            raise ValueError(f"BatchedMobility: unknown precision {precision!r}")
        eps = np.broadcast_to(np.asarray(eps, float), (len(systems),)).copy()
        pts, fines, w, pm, rot, pbody, nptr, ncol, nval = [], [], [], [], [], [], [np.zeros(1, np.int64)], [], []
        poff, foff, boff = [0], [0], [0]
        self.systems = list(systems)
        for sw in self.systems:
            points, fine, NN = sw.getSwimmerBnd()
            _, _, _, prot, frot = sw.getSwimmers()
            sizes, _ = sw.getSwimmerSizes()
            nsw = sw.swimmerNo()
            n, nb = len(points), len(prot)
            body = np.full(n, -1, np.int32)
            body[:nb] = np.repeat(np.arange(nsw, dtype=np.int32), sizes)
            r = np.zeros((n, 3))
            r[:nb] = prot
            if len(fine):
                NN = sp.csr_matrix(NN)
                NN.sort_indices()
                wj = np.asarray(NN.sum(axis=0)).ravel()
                frot_all = np.zeros((len(fine), 3))
                frot_all[:len(frot)] = frot
                arm = NN.T @ frot_all
                nptr.append(NN.indptr[1:].astype(np.int64) + nptr[-1][-1])
                ncol.append(NN.indices.astype(np.int32) + poff[-1])
                nval.append(NN.data)
            else:   # without quadrature the force points are their own sources
                wj, arm, fine = np.ones(n), r.copy(), points
                nptr.append(np.arange(1, n + 1, dtype=np.int64) + nptr[-1][-1])
                ncol.append(np.arange(n, dtype=np.int32) + poff[-1])
                nval.append(np.ones(n))
            pts.append(points)
            fines.append(fine)
            w.append(wj)
            pm.append(arm)
            rot.append(r)
            pbody.append(body)
            poff.append(poff[-1] + n)
            foff.append(foff[-1] + len(fine))
            boff.append(boff[-1] + nsw)
        cat = lambda a, d: np.ascontiguousarray(np.concatenate(a), dtype=d)  # noqa: E731
        self._arrs = [np.asarray(poff, np.int64), cat(pts, np.float64), np.asarray(foff, np.int64),
                      cat(fines, np.float64), np.ascontiguousarray(np.concatenate(nptr), np.int64),
                      cat(ncol, np.int32), cat(nval, np.float64), cat(w, np.float64), cat(pm, np.float64),
                      cat(rot, np.float64), cat(pbody, np.int32), np.asarray(boff, np.int64), eps]
        a = self._arrs
        h = ctypes.c_void_p()
        check(lib().nnk_batch_create(len(self.systems), ptr(a[0], _lib._i64), ptr(a[1]), ptr(a[2], _lib._i64),
                                     ptr(a[3]), ptr(a[4], _lib._i64), ptr(a[5], _lib._i32), ptr(a[6]),
                                     ptr(a[7]), ptr(a[8]), ptr(a[9]), ptr(a[10], _lib._i32),
                                     ptr(a[11], _lib._i64), ptr(a[12]), float(mu), _PREC[precision],
                                     ctypes.byref(h)))
        self.h = h
        if hipGraph:
            check(lib().nnk_batch_set_graph(h, 1))
        if autotune:
            check(lib().nnk_batch_set_autotune(h, 1))
        n = np.zeros(1, np.int64)
        check(lib().nnk_batch_size(h, ptr(n, _lib._i64)))
        self.n = int(n[0])
        self.npts = np.diff(a[0])
        self.nfine = np.diff(a[2])
        self.nsw = np.diff(a[11])
        self.offsets = np.concatenate([[0], np.cumsum(3 * self.npts + 6 * self.nsw)])
        self.last_solve = {}

    def p2pWaves(self):
        """Wavefronts per workgroup of the batched P2P (autotuned on the
        first operator application when autotune=True; 4 otherwise)."""
        v = ctypes.c_int32()
        check(lib().nnk_batch_p2p_waves(self.h, ctypes.byref(v)))
        return int(v.value)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().nnk_batch_destroy(self.h)
            except Exception:
                pass
            self.h = None

    def split(self, x):
        """The per-problem pieces [f (3N); U (3 nsw); Omega (3 nsw)] of a
        batched vector."""
        return [x[self.offsets[p]:self.offsets[p + 1]] for p in range(len(self.systems))]

    def rhs(self, fluidFunc=None, t=0.0):
        """[u_body - u_fluid; F; M] per problem (mobilityProblem.m:37-46)."""
        from .mobility import _rhs
        f = fluidFunc or (lambda t, x: np.zeros_like(x))
        return np.concatenate([_rhs(sw, f, t) for sw in self.systems])

    def apply(self, x, transpose=False):
        x = c64(x).ravel()
        y = np.empty_like(x)
        fn = lib().nnk_batch_apply_transpose if transpose else lib().nnk_batch_apply
        check(fn(self.h, ptr(x), ptr(y)))
        return y

    def solve(self, rhs, tol=1e-6, maxIter=200, transpose=False):
        """GMRES on every problem (A x = rhs, or A^T x = rhs); per-problem
        flag, relres, iter in last_solve."""
        b = c64(rhs).ravel()
        P = len(self.systems)
        x = np.empty_like(b)
        flag = np.zeros(P, np.int32)
        it = np.zeros(P, np.int32)
        rel = np.zeros(P)
        check(lib().nnk_batch_solve(self.h, ptr(b), float(tol), int(maxIter), int(bool(transpose)), ptr(x),
                                    ptr(flag, _lib._i32), ptr(rel), ptr(it, _lib._i32)))
        self.last_solve = {"flag": flag, "relres": rel, "iter": it}
        return x

    def gradient(self, z, lam):
        """dJ/d(force-point positions), dJ/d(quadrature positions), dJ/deps
        per problem, for A z = b and A^T lam = dJ/dz (b and the nearest map
        held fixed): lists of (N_p, 3) and (Q_p, 3) arrays and an array."""
        z, lam = c64(z).ravel(), c64(lam).ravel()
        gx = np.empty((int(self.npts.sum()), 3))
        gy = np.empty((max(int(self.nfine.sum()), 1), 3))
        ge = np.empty(len(self.systems))
        check(lib().nnk_batch_gradient(self.h, ptr(z), ptr(lam), ptr(gx), ptr(gy), ptr(ge)))
        po, fo = self._arrs[0], self._arrs[2]
        return ([gx[po[p]:po[p + 1]] for p in range(len(self.systems))],
                [gy[fo[p]:fo[p + 1]] for p in range(len(self.systems))], ge)
