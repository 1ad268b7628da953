"""Host-side point generators for the benchmark configurations.

These are this library's own generators (the reference's Geometries/ are
input scripts outside the accelerated path):
  cube_sphere       6 n^2 face-centred cube points projected to a sphere
                    (384 force points at n = 8, 6144 quadrature points at 32)
  prolate_spheroid  the same construction stretched along z (semi-axes
                    a, a, c)
  plane_wall        an n x n grid on a square in the plane z = z0
"""
import numpy as np


def cube_sphere(n, radius=1.0, centre=(0.0, 0.0, 0.0)):
    t = (np.arange(n) + 0.5) / n * 2 - 1
    a, b = np.meshgrid(t, t, indexing="ij")
    a, b = a.ravel(), b.ravel()
    faces = []
    for axis in range(3):
        for s in (-1.0, 1.0):
            p = np.empty((len(a), 3))
            p[:, axis] = s
            p[:, (axis + 1) % 3] = a
            p[:, (axis + 2) % 3] = b
            faces.append(p)
    p = np.concatenate(faces)
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    return radius * p + np.asarray(centre, float)


def prolate_spheroid(n, a, c, centre=(0.0, 0.0, 0.0)):
    p = cube_sphere(n)
    return p * np.array([a, a, c]) + np.asarray(centre, float)


def plane_wall(n, half_width, z0=0.0, centre=(0.0, 0.0)):
    g = (np.arange(n) + 0.5) / n * 2 * half_width - half_width
    X, Y = np.meshgrid(g + centre[0], g + centre[1], indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), np.full(X.size, float(z0))], 1)
