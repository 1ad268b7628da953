"""Time stepping with operator reuse (GPU).

The operator cache must not change results: a handle built from cached
per-level operators gives bit-identical velocities to one built cold.  The
time stepper around mobilityProblem is checked on a problem with a closed
form: a torque- and force-free sphere centred in the simple shear
u = (g z, 0, 0) rotates at half the vorticity, Omega = (0, g/2, 0), and does
not translate; its basis vectors b1, b2 follow the rotation about y by the
angle g t / 2.
"""
import gc

import numpy as np
import pytest

from geometry import cube_sphere

pytestmark = pytest.mark.gpu


def _fmm(nnk, p, prec):
    tree = nnk.OcTree(p, p, "nodeCapacity", 32, "rootCube", [0.0, 1.0])
    return nnk.KIFMM(tree, [1e-3, 1.0], "precision", prec)


@pytest.mark.parametrize("prec", ["fp64", "mixedh"])
def test_operator_cache_reuse_is_bit_identical(nnk, prec):
    rng = np.random.default_rng(3)
    p1, p2 = rng.random((6000, 3)), rng.random((6000, 3))
    q = rng.standard_normal((6000, 3))
    nnk.operatorCache(0)
    gc.collect()
    cold = _fmm(nnk, p2, prec).computeVel(q)
    gc.collect()
    nnk.operatorCache(16 << 30)
    f1 = _fmm(nnk, p1, prec)
    f1.computeVel(q)
    del f1
    gc.collect()
    hits0, _, kept, entries = nnk.operatorCacheStats()
    assert entries > 0 and kept > 0
    warm = _fmm(nnk, p2, prec).computeVel(q)
    assert nnk.operatorCacheStats()[0] > hits0
    np.testing.assert_array_equal(warm, cold)
    nnk.operatorCache(0)
    gc.collect()
    assert nnk.operatorCacheStats()[3] == 0


def test_sphere_in_shear_rotates_at_half_the_vorticity(nnk):
    sw = nnk.Swimmers()
    c, fine = cube_sphere(16), cube_sphere(64)
    NN = nnk.nearestMatrix(c, fine)
    sw.addSwimmer(c, fine, NN, [0, 0, 0], [1, 0, 0, 0, 1, 0], np.zeros_like(c), [0, 0, 0], [0, 0, 0])
    g = 1.0

    def shear(t, x):
        v = np.zeros_like(x)
        v[:, 0] = g * x[:, 2]
        return v
    tree_par = ["nodeCapacity", 64, "nearest", 1, "rootCube", [-1.5, 1.5]]

    def rhs(t, x):
        return nnk.mobilityProblem(sw, shear, t, x, [0.02, 1.0], tree_par, 0, 1, 0.1, 1e-8, 200, [1e-6, 100])
    y0 = np.r_[0.0, 0, 0, 1, 0, 0, 0, 1, 0]
    nnk.operatorCache(0)
    gc.collect()
    cold = rhs(0.0, y0)
    nnk.operatorCache(8 << 30)
    hits0 = nnk.operatorCacheStats()[0]
    warm0 = rhs(0.0, y0)
    warm = rhs(0.0, y0)
    assert nnk.operatorCacheStats()[0] > hits0
    np.testing.assert_array_equal(warm, cold)
    np.testing.assert_array_equal(warm0, cold)
    # one step's rates: Omega x b1 = (0, 0, -g/2), Omega x b2 = 0, U = 0
    assert np.abs(cold[0:3]).max() < 1e-6
    np.testing.assert_allclose(cold[3:6], [0, 0, -g / 2], atol=0.01 * g)
    assert np.abs(cold[6:9]).max() < 1e-6
    T = 0.6
    t, y = nnk.ode45(rhs, [0, T], y0)
    nnk.operatorCache(0)
    th = g / 2 * t
    ang = np.arctan2(-y[:, 5], y[:, 3])
    assert np.max(np.abs(ang - th)) < 0.02 * th[-1], (ang[-1], th[-1])
    np.testing.assert_allclose(np.linalg.norm(y[:, 3:6], axis=1), 1, atol=1e-4)
    assert np.abs(y[:, 0:3]).max() < 1e-5
    np.testing.assert_allclose(y[:, 6:9], np.tile([0, 1, 0], (len(t), 1)), atol=1e-5)
