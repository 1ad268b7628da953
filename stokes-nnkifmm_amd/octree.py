"""OcTree -- host mirror of the reference's Classes/OcTree.m.

Same constructor arguments (name/value pairs or keywords), same properties
and the same 1-based node numbering; the tree and its U/V/W/X lists are
built by libnnkifmm (nnk_tree_create).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import c64, check, lib, ptr

_OPTS = {"parThreads": 0, "nodeCapacity": 200, "maxDepth": 21, "finePoints": None, "nearest": 0,
         "NN": None, "blockSize": 0.2, "rootCube": None}


def parse_options(defaults, args, kwargs, where):
    """MATLAB inputParser semantics: name/value pairs after the required
    arguments, here also accepted as keywords."""
    out = dict(defaults)
    if len(args) % 2:
        raise ValueError(f"{where}: name/value arguments must come in pairs")
    for k, v in zip(args[0::2], args[1::2]):
        kwargs = dict(kwargs)
        kwargs[k] = v
    for k, v in kwargs.items():
        if k not in defaults:
            raise ValueError(f"{where}: unknown option '{k}'")
        out[k] = v
    return out


class Interactions:
    """interactions{node, kind} with node 1-based and kind 1..4 = U,V,W,X
    (Classes/OcTree.m:270-288); also interactions[node-1][kind-1]."""

    def __init__(self, ptrs, idxs):
        self._ptr = ptrs
        self._idx = idxs

    def get(self, node, kind):
        p = self._ptr[kind - 1]
        return self._idx[kind - 1][p[node - 1]:p[node]]

    def __call__(self, node, kind):
        return self.get(node, kind)

    def __getitem__(self, i):
        return [self.get(i + 1, k) for k in (1, 2, 3, 4)]

    def __len__(self):
        return len(self._ptr[0]) - 1


class OcTree:
    """OcTree(potPoints, targetPoints, 'nodeCapacity', c, 'maxDepth', d,
    'nearest', 0|1, 'finePoints', F, 'NN', NN, ...)   Classes/OcTree.m:77-129

    One option beyond the reference: 'rootCube', [lo, hi] fixes the root box
    instead of taking the points' bounding cube."""

    def __init__(self, potPoints, targetPoints, *args, **kwargs):
        a = parse_options(_OPTS, args, kwargs, "OcTree")
        pot = c64(potPoints).reshape(-1, 3)
        tgt = c64(targetPoints).reshape(-1, 3)
        h = ctypes.c_void_p()
        if a["rootCube"] is None:
            check(lib().nnk_tree_create(ptr(pot), len(pot), ptr(tgt), len(tgt), int(a["nodeCapacity"]),
                                        int(a["maxDepth"]), ctypes.byref(h)))
        else:  # extension: fixed root [lo lo lo hi hi hi] (time stepping reuses its operators)
            cube = c64(a["rootCube"]).ravel()
            if cube.size != 2:
                raise ValueError("OcTree: rootCube must be [lo, hi]")
            check(lib().nnk_tree_create_in_cube(ptr(pot), len(pot), ptr(tgt), len(tgt), int(a["nodeCapacity"]),
                                                int(a["maxDepth"]), ptr(cube), ctypes.byref(h)))
        self._h = h
        self.buildTimes = np.zeros(4)   # ms: de-duplication, subdivision, point lists, U/V/W/X lists
        check(lib().nnk_tree_build_times(h, ptr(self.buildTimes)))
        npts, npot, ntgt = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        cnt, cap = ctypes.c_int32(), ctypes.c_int32()
        check(lib().nnk_tree_sizes(h, ctypes.byref(npts), ctypes.byref(npot), ctypes.byref(ntgt),
                                   ctypes.byref(cnt), ctypes.byref(cap)))
        self.nodeCount = cnt.value
        self.nodeCapacity = cap.value
        n, m = cnt.value, npts.value
        self.points = np.empty((m, 3))
        pots = np.empty(m, np.uint8)
        tgts = np.empty(m, np.uint8)
        self.pointIndex = np.empty(m, np.int32)
        self.nodeCorners = np.empty((n, 6))
        self.nodeLevel = np.empty(n, np.int32)
        self.nodeParents = np.empty(n, np.int32)
        self.nodeChildren = np.empty((n, 8), np.int32)
        check(lib().nnk_tree_export(h, ptr(self.points), ptr(pots, _lib._u8), ptr(tgts, _lib._u8),
                                    ptr(self.pointIndex, _lib._i32), ptr(self.nodeCorners),
                                    ptr(self.nodeLevel, _lib._i32), ptr(self.nodeParents, _lib._i32),
                                    ptr(self.nodeChildren, _lib._i32)))
        self.potentials = pots.astype(bool)
        self.targets = tgts.astype(bool)
        ptrs, idxs = [], []
        for kind in range(4):
            tot = ctypes.c_int64()
            check(lib().nnk_tree_list_size(h, kind, ctypes.byref(tot)))
            p = np.zeros(n + 1, np.int64)
            ix = np.empty(max(tot.value, 1), np.int32)
            check(lib().nnk_tree_list_export(h, kind, ptr(p, _lib._i64), ptr(ix, _lib._i32)))
            ptrs.append(p)
            idxs.append(ix[:tot.value])
        self.interactions = Interactions(ptrs, idxs)
        self.finePoints = None
        self.NN = None
        fine = a["finePoints"]
        if a["nearest"] and fine is not None and len(fine):
            self.finePoints = c64(fine).reshape(-1, 3)
            a["finePoints"] = None
        if a["nearest"]:
            if a["NN"] is None or (hasattr(a["NN"], "shape") and 0 in a["NN"].shape):
                if self.finePoints is None:
                    raise ValueError("OcTree: nearest mode needs finePoints")
                from .nearest import nearestMatrix
                self.NN = nearestMatrix(pot, self.finePoints)
            else:
                self.NN = a["NN"]
            a["NN"] = None
            self._attach_nearest()
        else:
            self.NN = a["NN"]
            a["NN"] = None
        self.arguments = a

    def _attach_nearest(self):
        import scipy.sparse as sp
        NN = sp.csc_matrix(self.NN)
        NN.sort_indices()
        if NN.shape != (len(self.finePoints), int(self.potentials.sum())):
            raise ValueError(f"OcTree: NN must be {len(self.finePoints)} x {int(self.potentials.sum())}")
        self.NN = NN
        colptr = np.ascontiguousarray(NN.indptr, np.int64)
        row = np.ascontiguousarray(NN.indices, np.int64)
        val = c64(NN.data)
        check(lib().nnk_tree_set_nearest(self._h, ptr(self.finePoints), len(self.finePoints),
                                         ptr(colptr, _lib._i64), ptr(row, _lib._i64), ptr(val)))

    @property
    def handle(self):
        return self._h

    def getPotentialPoints(self):
        return self.points[self.potentials]

    def getTargetPoints(self):
        return self.points[self.targets]

    def getFinePoints(self):
        return self.finePoints

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                lib().nnk_tree_destroy(h)
            except Exception:
                pass
            self._h = None
