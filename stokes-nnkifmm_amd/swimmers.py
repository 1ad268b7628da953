"""Swimmers -- host mirror of the reference's Classes/Swimmers.m.

Bodies (force points in the body frame, optional fine quadrature points and
NN map, centre x0, basis b1/b2, point velocities, total force and moment)
followed by an optional fixed boundary.  Rotation/translation bookkeeping
is host-side numpy, as in the reference; genTree/genKIFMM build the GPU
tree and operator.
"""
import numpy as np

from .kifmm import KIFMM
from .octree import OcTree


def _blkdiag(mats):
    import scipy.sparse as sp
    mats = [m for m in mats if m is not None and np.prod(getattr(m, "shape", (0, 0))) > 0]
    if not mats:
        return sp.csc_matrix((0, 0))
    return sp.block_diag(mats, format="csc")


def _basis(b1, b2):
    b1 = np.asarray(b1, float).ravel()
    b2 = np.asarray(b2, float).ravel()
    return np.stack([b1, b2, np.cross(b1, b2)], 1)   # B = [b1 b2 b1xb2]


class Swimmers:
    def __init__(self):
        self.x0, self.b = [], []
        self.Geometries, self.Vel, self.swSize = [], [], []
        self.GeometriesFine, self.swSizeFine, self.NN = [], [], []
        self.F, self.Moment = [], []
        self.Boundary = np.zeros((0, 3))
        self.BoundaryVel = np.zeros((0, 3))
        self.BoundaryFine = np.zeros((0, 3))
        self.BoundaryNN = None
        self.Tree = None
        self.FMM = None

    def addSwimmer(self, *a):
        """(points, centre, basis, vel, F, Moment) or, with NEAREST,
        (points, finePoints, NN, centre, basis, vel, F, Moment)
        (Swimmers.m:83-125)."""
        if len(a) == 6:
            pts, x0, b, vel, F, M = a
            fine, NN = np.zeros((0, 3)), None
        elif len(a) == 8:
            pts, fine, NN, x0, b, vel, F, M = a
        else:
            raise ValueError("Invalid number of inputs")
        b = np.asarray(b, float).ravel()
        self.Geometries.append(np.asarray(pts, float).reshape(-1, 3))
        self.swSize.append(len(self.Geometries[-1]))
        self.GeometriesFine.append(np.asarray(fine, float).reshape(-1, 3))
        self.swSizeFine.append(len(self.GeometriesFine[-1]))
        self.NN.append(NN)
        self.x0.append(np.asarray(x0, float).ravel())
        self.b.append((b[0:3], b[3:6]))
        self.Vel.append(np.asarray(vel, float).reshape(-1, 3))
        self.F.append(np.asarray(F, float).ravel())
        self.Moment.append(np.asarray(M, float).ravel())

    def updateBnd(self, *a):
        """(points, vel) or (points, finePoints, NN, vel) (Swimmers.m:127-148)."""
        if len(a) == 2:
            self.Boundary = np.asarray(a[0], float).reshape(-1, 3)
            self.BoundaryFine, self.BoundaryNN = np.zeros((0, 3)), None
            self.BoundaryVel = np.asarray(a[1], float).reshape(-1, 3)
        elif len(a) == 4:
            self.Boundary = np.asarray(a[0], float).reshape(-1, 3)
            self.BoundaryFine = np.asarray(a[1], float).reshape(-1, 3)
            self.BoundaryNN = a[2]
            self.BoundaryVel = np.asarray(a[3], float).reshape(-1, 3)
        else:
            raise ValueError("Invalid number of inputs")

    def swimmerNo(self):
        return len(self.Geometries)

    def getSwimmerSizes(self):
        return list(self.swSize), list(self.swSizeFine)

    def _index(self, sizes):
        return np.concatenate([[0], np.cumsum(sizes)]).astype(int)

    def getSwimmers(self):
        """(points, finePoints, NN, pointsrot, finerot): rotated+translated and
        rotated-only coordinates (Swimmers.m:150-200)."""
        n = self.swimmerNo()
        pts = np.concatenate(self.Geometries) if n else np.zeros((0, 3))
        fine = np.concatenate(self.GeometriesFine) if n else np.zeros((0, 3))
        NN = _blkdiag(self.NN)
        si, sf = self._index(self.swSize), self._index(self.swSizeFine)
        prot, frot = np.zeros_like(pts), np.zeros_like(fine)
        pout, fout = pts.copy(), fine.copy()
        for i in range(n):
            B = _basis(*self.b[i])
            prot[si[i]:si[i + 1]] = pts[si[i]:si[i + 1]] @ B.T
            pout[si[i]:si[i + 1]] = prot[si[i]:si[i + 1]] + self.x0[i]
            if len(fine):
                frot[sf[i]:sf[i + 1]] = fine[sf[i]:sf[i + 1]] @ B.T
                fout[sf[i]:sf[i + 1]] = frot[sf[i]:sf[i + 1]] + self.x0[i]
        return pout, fout, NN, prot, frot

    def getIndSwimmer(self, index):
        """(points, finePoints, NN, x0) of swimmer index (1-based), rotated
        but not translated (Swimmers.m:202-230)."""
        i = index - 1
        B = _basis(*self.b[i])
        fine = self.GeometriesFine[i]
        return (self.Geometries[i] @ B.T, fine @ B.T if len(fine) else fine, self.NN[i], self.x0[i])

    def getBnd(self):
        return self.Boundary, self.BoundaryFine, self.BoundaryNN

    def getSwimmerBnd(self):
        """Swimmers first, then the boundary (Swimmers.m:254-310)."""
        p, f, NN, _, _ = self.getSwimmers()
        points = np.concatenate([p, self.Boundary])
        fine = np.concatenate([f, self.BoundaryFine])
        NN = _blkdiag([NN, self.BoundaryNN])
        return points, fine, NN

    def getVelocities(self):
        n = self.swimmerNo()
        out = []
        for i in range(n):
            out.append(self.Vel[i] @ _basis(*self.b[i]).T)
        out.append(self.BoundaryVel)
        return np.concatenate(out)

    def getForceMoment(self):
        n = self.swimmerNo()
        if not n:
            return np.zeros(0), np.zeros(0)
        return np.concatenate(self.F), np.concatenate(self.Moment)

    def updateSwimmmer(self, index, x0, b):
        b = np.asarray(b, float).ravel()
        self.x0[index - 1] = np.asarray(x0, float).ravel()
        self.b[index - 1] = (b[0:3], b[3:6])

    def updateSwimmmerRHS(self, index, G, offset):
        """Gravity: F = G, M = -(4/3) pi offset (b3 x G) (Swimmers.m:372-385)."""
        b1, b2 = self.b[index - 1]
        b3 = np.cross(b1, b2)
        G = np.asarray(G, float).ravel()
        self.F[index - 1] = G
        self.Moment[index - 1] = -4.0 / 3.0 * np.pi * offset * np.cross(b3, G)

    def genTree(self, *args, **kwargs):
        points, fine, NN = self.getSwimmerBnd()
        opts = dict(kwargs)
        for k, v in zip(args[0::2], args[1::2]):
            opts[k] = v
        opts.setdefault("finePoints", fine)
        opts.setdefault("NN", NN if NN.shape[0] else None)
        self.Tree = OcTree(points, points, **opts)
        return self

    def genKIFMM(self, kernelPar, *args, **kwargs):
        self.FMM = KIFMM(self.Tree, kernelPar, *args, **kwargs)
        self._mobility = None
        return self
