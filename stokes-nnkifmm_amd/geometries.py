"""Host-side point generators for the benchmark configurations.

These are this library's own generators (the reference's Geometries/ are
input scripts outside the accelerated path):
  cube_sphere       6 n^2 face-centred cube points projected to a sphere
                    (384 force points at n = 8, 6144 quadrature points at 32)
  prolate_spheroid  the same construction stretched along z (semi-axes
                    a, a, c)
  plane_wall        an n x n grid on a square in the plane z = z0
  random_frames     seeded non-overlapping centres and random orientations
  spheroid_suspension  bodies + wall assembled into a Swimmers system with
                    NEAREST quadrature (the adaptive mobility workload:
                    64 prolate spheroids, aspect 3, semi-major axis 1, in a
                    20^3 box above a 40 x 40 wall, 4x refined quadrature)
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


def random_frames(nbody, box, min_dist, zmin, seed):
    """Centres uniform in [-box/2+1, box/2-1]^2 x [zmin, box-1], at least
    min_dist apart (rejection sampling), and uniformly random orientations
    (QR of a Gaussian matrix); returns (centres, b1, b2)."""
    rng = np.random.default_rng(seed)
    lo = np.array([-box / 2 + 1, -box / 2 + 1, zmin])
    hi = np.array([box / 2 - 1, box / 2 - 1, box - 1])
    centres = []
    tries = 0
    while len(centres) < nbody:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("random_frames: cannot place the bodies")
        c = rng.uniform(lo, hi)
        if all(np.linalg.norm(c - d) >= min_dist for d in centres):
            centres.append(c)
    b1, b2 = [], []
    for _ in range(nbody):
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        q = q * np.sign(np.diag(r))
        if np.linalg.det(q) < 0:
            q[:, 2] = -q[:, 2]
        b1.append(q[:, 0])
        b2.append(q[:, 1])
    return np.array(centres), np.array(b1), np.array(b2)


def spheroid_suspension(nbody=64, n=18, refine=4, a=1.0 / 3.0, c=1.0, box=20.0, wall_n=40, seed=7,
                        force=(0.0, 0.0, -1.0)):
    """nbody prolate spheroids (semi-axes a, a, c; 6 n^2 force points,
    6 (refine n)^2 quadrature points with the nearest map) at random
    non-overlapping positions and orientations in a box^3 region above a
    wall_n x wall_n no-slip wall at z = 0 spanning the box (quadrature
    refine x finer).  Each body carries the force `force` and no moment.
    Returns a Swimmers system (tree and operator not built yet)."""
    from .nearest import nearestMatrix
    from .swimmers import Swimmers
    body = prolate_spheroid(n, a, c)
    fine = prolate_spheroid(refine * n, a, c)
    NN = nearestMatrix(body, fine)
    centres, b1, b2 = random_frames(nbody, box, 2.0 * c + 0.1, c + 0.2, seed)
    sw = Swimmers()
    for i in range(nbody):
        sw.addSwimmer(body, fine, NN, centres[i], np.concatenate([b1[i], b2[i]]), np.zeros_like(body),
                      np.asarray(force, float), np.zeros(3))
    if wall_n > 0:
        wall = plane_wall(wall_n, box / 2)
        wfine = plane_wall(refine * wall_n, box / 2)
        sw.updateBnd(wall, wfine, nearestMatrix(wall, wfine), np.zeros_like(wall))
    return sw
