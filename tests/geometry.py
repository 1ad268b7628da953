"""Point sets shared by the tests (plain numpy, no reference code)."""
import numpy as np


def cube_sphere(n, radius=1.0, centre=(0.0, 0.0, 0.0)):
    """6 n^2 points: face centres of an n x n grid on each cube face,
    projected to the sphere (the force-point layout of the one-sphere
    mobility configuration, 6*8^2 = 384 points for n = 8)."""
    t = (np.arange(n) + 0.5) / n * 2 - 1
    a, b = np.meshgrid(t, t, indexing="ij")
    a, b = a.ravel(), b.ravel()
    one = np.ones_like(a)
    faces = []
    for axis in range(3):
        for s in (-1.0, 1.0):
            p = np.empty((len(a), 3))
            p[:, axis] = s * one
            p[:, (axis + 1) % 3] = a
            p[:, (axis + 2) % 3] = b
            faces.append(p)
    p = np.concatenate(faces)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return radius * p + np.asarray(centre)


def grid16():
    """16^3 points x_i = i/15: with nodeCapacity 1 this gives the complete
    octree of depth 4 over [0,1]^3 (one point per level-4 cell)."""
    x = np.arange(16) / 15.0
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def blobs(rng):
    return np.concatenate([rng.normal(0.3, 0.02, (1500, 3)), rng.random((1500, 3)),
                           rng.normal(0.7, 0.05, (1000, 3))])
