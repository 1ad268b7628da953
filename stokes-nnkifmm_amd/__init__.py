"""MI355X-native KIFMM matvec and mobility solve for regularized Stokeslets.

Host mirror of the reference's MATLAB operator API (Classes/OcTree.m,
Classes/KIFMM.m) over the C ABI of libnnkifmm (include/nnkifmm.h).
Import as ``importlib.import_module("stokes-nnkifmm_amd")``.
"""
from ._lib import NNKError, lib  # noqa: F401
from .batch import BatchedMobility  # noqa: F401
from .kifmm import KIFMM  # noqa: F401
from .mobility import (genMatrixComponets, mobility_gmres, mobility_refined, mobilityMatrixMulti,  # noqa: F401
                       mobilityPrecon,
                       mobilityProblem, mobilityProblemGravity)
from .nearest import nearestMatrix  # noqa: F401
from .resistance import Gen_Resistance, Rigid_Resistance, resistanceSolve, velocityModes  # noqa: F401
from .octree import OcTree  # noqa: F401
from .swimmers import Swimmers  # noqa: F401
from .timestepping import (ode45, odeset, operatorCache, operatorCacheStats, trajectoryOutput,  # noqa: F401
                           writeTrajectory)


def library_path():
    from ._lib import LIBPATH
    return LIBPATH
