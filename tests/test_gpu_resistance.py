"""Resistance problems and the flow-field pipeline on the KIFMM operator
(GPU): the callers of gmres(@(x) fmm.computeVel(x), ...) in the reference
(Rigid_Resistance.m, Gen_Resistance.m, FlowField.m).

Pinned by: a dense FP64 solve of the same system (oracle kernel matrix,
the tree being the root and 8 leaves so the FMM is the direct sum); the
grand resistance matrix of a sphere, 6 pi mu a I and 8 pi mu a^3 I
(closed form, to discretisation accuracy); the flow field of a squirmer
evaluated by a second KIFMM with targets apart from the sources, against
the oracle's direct sum."""
import numpy as np
import pytest

from geometry import cube_sphere

pytestmark = pytest.mark.gpu


def _sphere(nnk, n=8, refine=4, eps=0.01, fmt=1):
    c, fine = cube_sphere(n), cube_sphere(refine * n)
    NN = nnk.nearestMatrix(c, fine)
    t = nnk.OcTree(c, c, "nodeCapacity", 64, "nearest", 1, "finePoints", fine, "NN", NN)
    return nnk.KIFMM(t, [eps, 1.0], "GMRES", 1, "format", fmt), c, fine, NN


@pytest.mark.parametrize("precon", [False, True])
def test_resistance_solve_matches_dense(oracle, nnk, precon):
    from importlib import import_module
    res = import_module("stokes-nnkifmm_amd.resistance")
    fmm, c, fine, NN = _sphere(nnk, n=4, refine=4)
    u = res.velocityModes(c, [0.3, -1.0, 0.5], [0.2, 0.0, -0.4])
    f = res.resistanceSolve(fmm, u.ravel(), 1e-12, 300, precon)
    assert res.last_solve["flag"] == 0
    K = oracle.kernel(fine.ravel(), c.ravel(), 0.01)              # 3N x 3Q, interleaved
    S = K @ np.kron(NN.toarray(), np.eye(3))
    fd = np.linalg.solve(S, u.ravel())
    assert np.linalg.norm(f - fd) <= 1e-8 * np.linalg.norm(fd)


def test_sphere_grand_resistance(nnk):
    from importlib import import_module
    res = import_module("stokes-nnkifmm_amd.resistance")
    fmm, *_ = _sphere(nnk, n=8, refine=4)
    R = res.Gen_Resistance(fmm, 1e-8, 300)
    ref = np.diag([6 * np.pi] * 3 + [8 * np.pi] * 3)
    assert np.abs(R - ref).max() <= 0.03 * 8 * np.pi, np.round(R, 3)
    assert np.abs(R[:3, 3:]).max() < 1e-3 * 6 * np.pi and np.abs(R[3:, :3]).max() < 1e-3 * 6 * np.pi


def test_flow_field_pipeline(oracle, nnk):
    """FlowField.m: resistance solve for a squirming surface, quadrature
    strengths (NN (x) I3) f, velocities on an evaluation grid from a KIFMM
    whose targets are not its sources."""
    from importlib import import_module
    res = import_module("stokes-nnkifmm_amd.resistance")
    geo = import_module("stokes-nnkifmm_amd.geometries")
    fmm, c, fine, NN = _sphere(nnk, n=6, refine=4, fmt=1)
    vel = geo.squirmer_slip(c, 2.0, 0.0)
    pot = res.resistanceSolve(fmm, vel.ravel(), 1e-8, 100).reshape(-1, 3)
    q = NN @ pot
    g = np.linspace(-3, 3, 14)
    X, Y, Z = np.meshgrid(g, g, g)
    ev = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    u = nnk.KIFMM(nnk.OcTree(fine, ev), [0.01, 1.0]).computeVel(q)
    ud = oracle.direct(fine, q, ev, 0.01)
    assert np.linalg.norm(u - ud) <= 1e-4 * np.linalg.norm(ud)
