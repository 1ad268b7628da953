"""Point sets shared by the tests."""
import importlib

import numpy as np

_geo = importlib.import_module("stokes-nnkifmm_amd.geometries")
cube_sphere = _geo.cube_sphere


def grid16():
    """16^3 points x_i = i/15: with nodeCapacity 1 this gives the complete
    octree of depth 4 over [0,1]^3 (one point per level-4 cell)."""
    x = np.arange(16) / 15.0
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def blobs(rng):
    return np.concatenate([rng.normal(0.3, 0.02, (1500, 3)), rng.random((1500, 3)),
                           rng.normal(0.7, 0.05, (1000, 3))])
