"""The reference's geometry generators mirrored in geometries.py (CPU):
Geometries/wall.m, prolateSpheroid.m, Rotate.m and unique(...,'rows',
'stable').  Pinned by their closed forms: grid sizes ceil(scale/h) and
meshgrid column-major order, the ring counts of prolateSpheroid, rotation
matrices orthogonal with determinant 1, first-occurrence order."""
import importlib
import math

import numpy as np

geo = importlib.import_module("stokes-nnkifmm_amd.geometries")


def test_wall_grid_and_order():
    w = geo.wall([5, 0, 0], [10, 10], 1, 0.3)
    assert w.shape == (34 * 34, 3)
    np.testing.assert_array_equal(w[:, 0], 5.0)
    # meshgrid(y, z) flattened column-major: z runs fastest
    np.testing.assert_allclose(w[:3, 1], [-5, -5, -5])
    np.testing.assert_allclose(w[:3, 2], [-5, -5 + 10 / 33, -5 + 20 / 33])
    np.testing.assert_allclose(w[-1], [5, 5, 5])
    z = geo.wall([0, 0, -2], [4, 6], 3, 1.0)
    assert z.shape == (4 * 6, 3) and np.all(z[:, 2] == -2)
    np.testing.assert_allclose([z[:, 0].min(), z[:, 0].max(), z[:, 1].min(), z[:, 1].max()], [-2, 2, -3, 3])


def test_prolate_spheroid_rings():
    a, c, h = 0.5, 1.5, 0.2
    p = geo.prolateSpheroid(a, c, h)
    vN = math.ceil(math.pi * (3 * (a + c) - math.sqrt((3 * a + c) * (a + 3 * c))) / (2 * h))
    v = np.linspace(0, np.pi, vN)
    expect = 2 + sum(math.ceil(2 * math.pi * a * math.sqrt(1 - math.cos(t) ** 2) / h) for t in v[1:-1])
    assert len(p) == expect
    np.testing.assert_allclose(p[:2], [[0, 0, c], [0, 0, -c]])
    np.testing.assert_allclose((p[:, 0] ** 2 + p[:, 1] ** 2) / a ** 2 + p[:, 2] ** 2 / c ** 2, 1, atol=1e-12)
    assert len(geo.prolateSpheroid(0.5, 0.5, 0.25)) == 51


def test_rotate_and_unique_rows():
    R = geo.Rotate(np.eye(3), 0.3, -1.1, 2.0, [0, 0, 0])
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
    assert abs(np.linalg.det(R) - 1) < 1e-14
    c = np.array([1.0, 2.0, 3.0])
    p = np.random.default_rng(0).random((10, 3))
    q = geo.Rotate(p, 0.3, -1.1, 2.0, c)
    np.testing.assert_allclose(np.linalg.norm(q - c, axis=1), np.linalg.norm(p - c, axis=1), atol=1e-14)
    rows = np.array([[1.0, 0, 0], [0, 1, 0], [1, 0, 0], [2, 2, 2], [0, 1, 0]])
    np.testing.assert_array_equal(geo.unique_rows(rows), rows[[0, 1, 3]])


def test_squirmer_slip_matches_the_reference_formula():
    # squirmerVel of Squiremers.m: scale sin(theta) e_theta, with e_theta
    # written as [x z, y z, -(x^2 + y^2)] / (|r| sqrt(x^2 + y^2))
    p = geo.prolateSpheroid(0.5, 0.5, 0.13)
    rho = np.sqrt(p[:, 0] ** 2 + p[:, 1] ** 2)
    r = np.linalg.norm(p, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        t = np.stack([p[:, 0] * p[:, 2], p[:, 1] * p[:, 2], -(rho ** 2)], 1) / (r * rho)[:, None]
    t[np.isnan(t)] = 0
    ref = 2.0 * np.sin(np.arccos(p[:, 2] / r))[:, None] * t
    np.testing.assert_allclose(geo.squirmer_slip(p, B1=2.0), ref, atol=1e-12)
