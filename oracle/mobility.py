"""Mobility-layer oracle -- TEST INFRASTRUCTURE ONLY (numpy + the C oracle).

Restates Mobility/MobilityMatrixMulti.m:17-36, MobilityPrecon.m:17-39 and
the gmres call of mobilityProblem.m:47-58 for small problems.  B and BT are
applied matrix-free from their definitions (genMatrixComponets.m:34-97):
  (B f)_n   = [ sum_{j in n} w_j f_j ;  sum_{j in n} p_j x f_j ]
  (BT y)_j  = -U_n - Omega_n x r_j      (j a point of body n; 0 on the boundary)
with r_j the rotated, untranslated point, w_j = 1 and p_j = r_j, or in
NEAREST mode w_j = sum_q NN(q,j) and p_j = sum_q NN(q,j) rfine_q.
Vectors: [f (3N: all x, all y, all z); U (3 nsw); Omega (3 nsw)].
GMRES: Saad & Schultz (1986) with modified Gram-Schmidt, x0 = 0.
Parity pins: Stokes drag on a sphere (tests/test_gpu_mobility.py), numpy
solves of small dense systems, and precon / apply / A1 / A2 against dense
matrices written from genMatrixComponets.m, ReducedSolve.m and
MobilityPrecon.m on the one-sphere tree (tests/test_oracle_mobility.py).
"""
import numpy as np


class MobilityOracle:
    def __init__(self, tree, eps, mu, body_ptr, r, w, p, res=6):
        self.t, self.eps, self.mu, self.res = tree, eps, mu, res
        self.ptr = np.asarray(body_ptr, int)
        self.nsw = len(self.ptr) - 1
        self.N = tree.npot
        self.r, self.w, self.p = np.asarray(r, float), np.asarray(w, float), np.asarray(p, float)
        self.n3 = 3 * self.N
        self.n = self.n3 + 6 * self.nsw
        A1 = np.zeros((6 * self.nsw, 6 * self.nsw))
        for k in range(6 * self.nsw):
            e = np.zeros(6 * self.nsw)
            e[k] = 1
            A1[:, k] = self.B(self.Bt(e))
        self.A1 = A1
        A2 = np.zeros_like(A1)
        for k in range(6 * self.nsw):
            e = np.zeros(6 * self.nsw)
            e[k] = 1
            A2[:, k] = self.BTt(self.BT(e))
        self.A2 = A2

    # --- B, B^T, BT, BT^T from their definitions ---
    def _split(self, y):
        s = self.nsw
        return y[0:3 * s].reshape(3, s).T, y[3 * s:6 * s].reshape(3, s).T

    def B(self, f):
        F = f.reshape(3, self.N).T
        out = np.zeros((2, self.nsw, 3))
        for n in range(self.nsw):
            a, b = self.ptr[n], self.ptr[n + 1]
            out[0, n] = (self.w[a:b, None] * F[a:b]).sum(0)
            out[1, n] = np.cross(self.p[a:b], F[a:b]).sum(0)
        return np.concatenate([out[0].T.ravel(), out[1].T.ravel()])

    def Bt(self, y):   # B^T y
        Fy, My = self._split(y)
        F = np.zeros((self.N, 3))
        for n in range(self.nsw):
            a, b = self.ptr[n], self.ptr[n + 1]
            F[a:b] = self.w[a:b, None] * Fy[n] + np.cross(My[n], self.p[a:b])
        return F.T.ravel()

    def BT(self, y):
        U, Om = self._split(y)
        V = np.zeros((self.N, 3))
        for n in range(self.nsw):
            a, b = self.ptr[n], self.ptr[n + 1]
            V[a:b] = -U[n] - np.cross(Om[n], self.r[a:b])
        return V.T.ravel()

    def BTt(self, v):  # BT^T v
        V = v.reshape(3, self.N).T
        out = np.zeros((2, self.nsw, 3))
        for n in range(self.nsw):
            a, b = self.ptr[n], self.ptr[n + 1]
            out[0, n] = -V[a:b].sum(0)
            out[1, n] = -np.cross(self.r[a:b], V[a:b]).sum(0)
        return np.concatenate([out[0].T.ravel(), out[1].T.ravel()])

    # --- operators ---
    def S(self, f, res=None):
        u = self.t.matvec(f.reshape(3, self.N).T, self.eps, self.mu, res or self.res, 1)
        return u.T.ravel()

    def R_inv(self, v):
        return self.t.reduced_solve(v.reshape(3, self.N).T, self.eps, self.mu).T.ravel()

    def apply(self, x):
        f, y = x[:self.n3], x[self.n3:]
        return np.concatenate([self.S(f) + self.BT(y), self.B(f)])

    def precon(self, x):
        f, uo = x[:self.n3], x[self.n3:]
        bf = self.R_inv(f)
        out = -np.linalg.solve(self.A2, self.BTt(self.S(self.Bt(np.linalg.solve(self.A1, uo)), 4)))
        bf = bf - self.R_inv(self.BT(out))
        return np.concatenate([bf, out])

    def gmres(self, b, tol, maxit, precon=False):
        op = (lambda v: self.apply(self.precon(v))) if precon else self.apply
        x, info = gmres(op, b, tol, maxit)
        if precon:
            x = self.precon(x)
        return x, info


def gmres(op, b, tol, maxit):
    """Unrestarted GMRES from x0 = 0 with modified Gram-Schmidt."""
    n = len(b)
    beta = np.linalg.norm(b)
    if beta == 0:
        return np.zeros(n), {"flag": 0, "iter": 0, "relres": 0.0}
    m = min(maxit, n)
    V = np.zeros((m + 1, n))
    H = np.zeros((m + 1, m))
    V[0] = b / beta
    k = 0
    for k in range(m):
        w = op(V[k])
        for j in range(k + 1):
            H[j, k] = V[j] @ w
            w = w - H[j, k] * V[j]
        H[k + 1, k] = np.linalg.norm(w)
        if H[k + 1, k] > 0:
            V[k + 1] = w / H[k + 1, k]
        e1 = np.zeros(k + 2)
        e1[0] = beta
        y, *_ = np.linalg.lstsq(H[:k + 2, :k + 1], e1, rcond=None)
        res = np.linalg.norm(H[:k + 2, :k + 1] @ y - e1)
        if res <= tol * beta or H[k + 1, k] == 0:
            break
    x = V[:k + 1].T @ y
    rel = np.linalg.norm(b - op(x)) / beta
    return x, {"flag": int(rel > tol), "iter": k + 1, "relres": rel}
