"""KIFMM -- host mirror of the reference's Classes/KIFMM.m.

The matvec runs in libnnkifmm on the GPU.  Inputs may be numpy arrays
(host; copied in and out) or torch CUDA tensors (device-resident; run on
torch's current stream).  Vector layouts follow the reference:
  GMRES = 0             potentials (N,3) -> velocities (M,3)
  GMRES = 1, format = 1 [x1 y1 z1 x2 ...]'      (KIFMM.m:84-85, 98-99)
  GMRES = 1, format = 2 [x1 .. xN y1 .. z1 ..]' (KIFMM.m:86-87, 100-101)
"""
import ctypes

import numpy as np

from ._lib import c64, check, lib, ptr
from .octree import parse_options

_OPTS = {"GPU": 0, "parThreads": 0, "blockSize": 0.1, "GMRES": 0, "coronaRes": 6, "coronaShells": 1,
         "format": 1, "precision": "fp64", "svdTol": 1e-8, "hipGraph": 0, "autotune": 0}
_PREC = {"fp64": 0, "mixed": 1, "mixed3": 2, "mixedh": 3, "fp64t": 4}


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class _Handle:
    def __init__(self, tree, eps, mu, res, shells, precision, svd_tol=1e-8, graph=False, autotune=False):
        h = ctypes.c_void_p()
        check(lib().nnk_fmm_create(tree.handle, float(eps), float(mu), int(res), int(shells),
                                   _PREC[precision], float(svd_tol), ctypes.byref(h)))
        self.h = h
        if graph:
            check(lib().nnk_fmm_set_graph(h, 1))
        if autotune:
            check(lib().nnk_fmm_set_autotune(h, 1))
        self.tree = tree  # the tree must outlive the handle

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().nnk_fmm_destroy(self.h)
            except Exception:
                pass
            self.h = None


class KIFMM:
    """KIFMM(tree, [eps mu], 'GPU', g, 'parThreads', p, 'blockSize', b,
    'GMRES', 0|1, 'coronaRes', r, 'coronaShells', s, 'format', 1|2)
    Classes/KIFMM.m:46-61.  'GPU', 'parThreads' and 'blockSize' are
    accepted for compatibility; the matvec always runs on the GPU and never
    materialises a kernel block.  Build-only options: 'precision' ('fp64',
    the reference's numbers, 'mixed' or 'mixed3': two- or three-part
    split-BF16 M2L, 'mixedh': two-part split-FP16 M2L on scaled operands,
    'fp64t': the mixed forms' truncated operators with every operation in
    FP64), 'svdTol' (the truncated forms' pseudo-inverse cut) and
    'hipGraph' (1: replay each matvec, and the mobility solve's GMRES
    operator, from a captured HIP graph) and 'autotune' (1: pick the M2L
    launch shapes by timing candidates on the first matvec)."""

    def __init__(self, tree, kernelPar, *args, **kwargs):
        self.tree = tree
        self.arguments = parse_options(_OPTS, args, kwargs, "KIFMM")
        if self.arguments["precision"] not in _PREC:
            raise ValueError(f"KIFMM: unknown precision {self.arguments['precision']!r}")
        kp = np.asarray(kernelPar, dtype=np.float64).ravel()
        if kp.size != 2:
            raise ValueError("KIFMM: kernelPar must be [eps mu]")
        self.arguments["kernelPar"] = kp
        self._handles = {}

    # -- device handles, one per corona resolution ------------------------
    def _handle(self, res=None, shells=None):
        a = self.arguments
        res = a["coronaRes"] if res is None else res
        shells = a["coronaShells"] if shells is None else shells
        key = (int(res), int(shells), a["precision"], float(a["kernelPar"][0]), float(a["kernelPar"][1]),
               float(a["svdTol"]), bool(a["hipGraph"]), bool(a["autotune"]))
        h = self._handles.get(key)
        if h is None:
            h = _Handle(self.tree, a["kernelPar"][0], a["kernelPar"][1], res, shells, a["precision"],
                        a["svdTol"], bool(a["hipGraph"]), bool(a["autotune"]))
            self._handles[key] = h
        return h

    # -- layouts ----------------------------------------------------------
    def _to_rows(self, x):
        a = self.arguments
        if a["GMRES"] == 1:
            if a["format"] == 1:
                return x.reshape(-1, 3)
            if a["format"] == 2:
                return x.reshape(3, -1).T
        return x.reshape(-1, 3)

    def _from_rows(self, u, like):
        a = self.arguments
        if a["GMRES"] == 1:
            if a["format"] == 1:
                return u.reshape(-1)
            if a["format"] == 2:
                return u.T.reshape(-1)
        return u

    def _run(self, potentials, h, reduced=False, solve=False):
        if _is_torch(potentials):
            import torch
            f = self._to_rows(potentials).to(torch.float64).contiguous()
            u = torch.empty((self.tree.targets.sum(), 3), dtype=torch.float64, device=f.device)
            stream = ctypes.c_void_p(torch.cuda.current_stream(f.device).cuda_stream)
            fn = lib().nnk_fmm_reduced_apply_device if reduced else lib().nnk_fmm_apply_device
            if solve:
                fn = lib().nnk_fmm_reduced_solve_device
            check(fn(h.h, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(u.data_ptr()), stream))
            return self._from_rows(u, potentials).contiguous()
        f = c64(self._to_rows(np.asarray(potentials)))
        u = np.empty((int(self.tree.targets.sum()), 3))
        fn = lib().nnk_fmm_reduced_apply if reduced else lib().nnk_fmm_apply
        if solve:
            fn = lib().nnk_fmm_reduced_solve
        check(fn(h.h, ptr(f), ptr(u)))
        return np.ascontiguousarray(self._from_rows(u, potentials))

    def computeVel(self, potentials):
        """KIFMM.m:63-105: upwardPass + downwardPass."""
        return self._run(potentials, self._handle())

    def computeVelSparse(self, potentials):
        """KIFMM.m:107-155: the same matvec with 4 points per corona edge."""
        return self._run(potentials, self._handle(res=4))

    def reducedComputeVel(self, potentials):
        """KIFMM.m:157-198: leaf-local product (reducedPass.m)."""
        return self._run(potentials, self._handle(), reduced=True)

    def reducedSolve(self, vel):
        """KIFMM.m:200-242: per-leaf block solve (Mobility/ReducedSolve.m)."""
        return self._run(vel, self._handle(), solve=True)
