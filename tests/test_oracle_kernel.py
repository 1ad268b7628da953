"""Oracle pins for the kernel, surfaces, LU and direct sum (CPU).

Each check comes from the mathematics, not from the oracle's own formula:
the regularized Stokeslet (Cortez, Fauci & Medovikov 2005, cited at
Tools/kernel.m:4) is symmetric, equals I/(4 pi mu eps) at r = 0 and tends
to the Oseen tensor (I + r r^T / r^2) / (8 pi mu r) for r >> eps.
"""
import numpy as np
import pytest


def test_kernel_transpose_symmetry(oracle):
    rng = np.random.default_rng(0)
    x, y = rng.random((7, 3)), rng.random((5, 3))
    A = oracle.kernel(x, y, 0.05, 1.3)   # K(x -> y): rows y
    B = oracle.kernel(y, x, 0.05, 1.3)
    assert A.shape == (15, 21)
    np.testing.assert_allclose(A, B.T, rtol=0, atol=1e-15 * np.abs(A).max())
    # each 3x3 block is symmetric
    blk = A.reshape(5, 3, 7, 3).transpose(0, 2, 1, 3)
    np.testing.assert_allclose(blk, blk.transpose(0, 1, 3, 2), rtol=0, atol=1e-15 * np.abs(A).max())


@pytest.mark.parametrize("eps,mu", [(1e-3, 1.0), (0.01, 2.5), (0.2, 0.7)])
def test_kernel_self_term(oracle, eps, mu):
    p = np.array([[0.3, -0.2, 1.1]])
    S = oracle.kernel(p, p, eps, mu)
    np.testing.assert_allclose(S, np.eye(3) / (4 * np.pi * mu * eps), rtol=1e-14)


def test_kernel_oseen_limit(oracle):
    rng = np.random.default_rng(1)
    r = rng.standard_normal(3)
    r *= 2.0 / np.linalg.norm(r)
    mu, eps = 0.9, 1e-7
    S = oracle.kernel(np.zeros((1, 3)), r[None], eps, mu)
    rn = np.linalg.norm(r)
    oseen = (np.eye(3) + np.outer(r, r) / rn**2) / (8 * np.pi * mu * rn)
    np.testing.assert_allclose(S, oseen, rtol=1e-12)
    # decay 1/r: doubling r halves the tensor in the Oseen regime
    S2 = oracle.kernel(np.zeros((1, 3)), 2 * r[None], eps, mu)
    np.testing.assert_allclose(S2, S / 2, rtol=1e-12)


def test_direct_sum_equals_dense_product(oracle):
    rng = np.random.default_rng(2)
    src, tgt, q = rng.random((40, 3)), rng.random((13, 3)), rng.standard_normal((40, 3))
    u = oracle.direct(src, q, tgt, 0.02, 1.1)
    S = oracle.kernel(src, tgt, 0.02, 1.1)
    np.testing.assert_allclose(u.ravel(), S @ q.ravel(), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("res,shells", [(4, 1), (6, 1), (8, 1), (6, 2)])
def test_surface_counts_and_coordinates(oracle, res, shells):
    c = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    up = oracle.surface(c, res, shells, False)
    assert len(up) == res**3 - (res - 2 * shells) ** 3
    g = np.round(up * (res - 1)).astype(int)
    np.testing.assert_allclose(up * (res - 1), g, atol=1e-12)
    inner = (g >= shells) & (g <= res - 1 - shells)
    assert not np.any(inner.all(axis=1))          # interior removed
    assert len({tuple(r) for r in g}) == len(g)   # no repeats
    # meshgrid(x,y,z) flattened column-major: y fastest, then x, then z
    key = g[:, 2] * res * res + g[:, 0] * res + g[:, 1]
    assert np.all(np.diff(key) > 0)
    dn = oracle.surface(c, res, shells, True)
    np.testing.assert_allclose(dn, 3 * up - 1, atol=1e-14)


def test_lu_solve_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 40))
    b = rng.standard_normal(40)
    np.testing.assert_allclose(oracle.lu_solve(A, b), np.linalg.solve(A, b), rtol=1e-10)
