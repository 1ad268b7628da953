"""Two small adaptive octrees whose W and X lists were worked out by hand
from the rules of Classes/OcTree.m (no code computed the expected lists).

Rules used (ids 1-based, nodeCapacity 1, so every box holding two or more
points splits):

* Child k (0-based) of a box is 4[x > c] + 2[y > c] + [z > c] (strict >,
  Dividenode, OcTree.m:199-213); the children of the box split first get
  ids count+1 .. count+8 and are split depth first (Divide, OcTree.m:165-178).
* U(i) = potential leaves whose closed box touches i (OcTree.m:323-327).
* W(i), i a target leaf (OcTree.m:343-360): nodes of level <= level(i)
  touching i, minus every parent of one of them; all their descendants
  (getChildren); keep potential leaves; setdiff with U (sorted).  Then, for
  level = max .. 1 and the nodes of that level in id order, a node whose 8
  children are all in W replaces them: W = [setdiff(W, children); node],
  i.e. the rest is re-sorted and the merged node is appended last.
* X(j) = find(Wpad == j) over the zero-padded matrix whose row r is W(r)
  (OcTree.m:279-288): column-major, so rows are ordered by the position j
  holds in their W first, then by row id.

Tree A (21 points).  Root [0,1]^3; its low octant P = node 2 = [0,.5]^3 splits
into 10..17 (10 = [0,.25]^3 holds the origin, 16 = [.25,.5]x[.25,.5]x[0,.25]
holds one point).  Node 14 = [.25,.5]x[0,.25]x[0,.25] splits into 18..25
(centre (.375,.125,.125)): 22 = [.375,.5]x[0,.125]x[0,.125] and
23 = [.375,.5]x[0,.125]x[.125,.25] hold one point per octant (8 each) and
split into 26..33 and 34..41; 24 and 25 hold one point each; 18..21 are
empty.  (1,1,1) sits alone in node 9.

  W(10): neighbours of level <= 2 are 1, 2 and 10..17 (all touch P's centre);
  1 and 2 are parents -> {10..17}.  Potential leaves below: 10, 16, 24, 25,
  26..41.  U(10) = [10 16] -> W = {24 25 26..41}.  Level 3, node 22:
  W = [24 25 34..41 22]; node 23: W = [sort(24 25 22) 23] = [22 24 25 23].
  Level 2: node 14's child 18 is not in W -> no merge.
  W(16): U(16) = [10 16 24 25] (24 and 25 touch the plane y = .25), so W
  starts as {26..41} -> 22 merged -> [34..41 22] -> 23 merged -> [22 23].
  W(9): only 9 touches 9 -> U = [9]; W starts {10 16 24 25 26..41} ->
  [10 16 22 24 25 23].
  W(24) and W(25): 22 and 23 each have children touching 24 (25), so no
  merge; W(24) = children of 22 and 23 not touching 24 (a child of 22
  touches 24 iff its y bit is 1, one of 23 iff y = 1 and z = 0).
  Level-4 leaves: every touching box is a leaf or a parent of a touching
  box, so getChildren returns [] and W = [].

  X(23): 23 is at position 1 of W(16), 3 of W(10), 5 of W(9) -> [16 10 9].
  X(22): position 0 of W(10) and W(16), 2 of W(9) -> [10 16 9].

Tree B (18 points).  Same skeleton, but 16 and 23 are empty and inside 22 the
child 26 = [.375,.4375]x[0,.0625]x[0,.0625] holds 8 points (one per octant,
split into 34..41) while 27..33 hold one point each; 24 holds one point.

  W(10): U(10) = [10]; W starts {24 27..41}.  Level 4 comes first: node 26
  (children 34..41) -> [24 27..33 26].  Then level 3: node 22 (children
  26..33, all now in W) -> [24 22].  (Top-down order would miss both.)
  W(9) = [10 24 22].  W(24): 22's children 28 29 32 33 touch 24 (y bit 1),
  so W starts {27 30 31 34..41} -> 26 merged -> [27 30 31 26].
  W(27..33): child c of 26 touches sibling s of 26 iff bits(c) contain
  bits(s) (s lies on 26's high faces in its set bits), so W(s) = the other
  children of 26, ascending.
"""
import numpy as np


def _octants(lo, hi):
    """One point per octant of a box, in child order: coordinate d is lo[d]
    or hi[d] (one either side of the box's centre)."""
    return [[x, y, z] for x in (lo[0], hi[0]) for y in (lo[1], hi[1]) for z in (lo[2], hi[2])]


def tree_a():
    pts = [[0, 0, 0], [1, 1, 1], [0.3, 0.3, 0.1], [0.45, 0.2, 0.05], [0.45, 0.2, 0.2]]
    pts += _octants((0.40, 0.03, 0.03), (0.47, 0.09, 0.09))   # node 22 -> 26..33
    pts += _octants((0.40, 0.03, 0.15), (0.47, 0.09, 0.22))   # node 23 -> 34..41
    return np.array(pts, float)


def tree_b():
    pts = [[0, 0, 0], [1, 1, 1], [0.45, 0.2, 0.05]]
    pts += _octants((0.39, 0.01, 0.01), (0.42, 0.05, 0.05))   # node 26 -> 34..41
    for k in range(1, 8):                                    # nodes 27..33
        pts.append([(0.40, 0.47)[k >> 2], (0.03, 0.09)[(k >> 1) & 1], (0.03, 0.09)[k & 1]])
    return np.array(pts, float)


def _w_b_leaf(s):
    """W of sibling s (27..33) of node 26 in tree B."""
    k = s - 26
    return [34 + c for c in range(8) if (c & k) != k]


# expected lists, ids 1-based; nodes not listed have empty W / X
TREE_A = {
    "nodeCount": 41,
    "U": {9: [9], 10: [10, 16], 16: [10, 16, 24, 25],
          24: [16, 24, 25, 28, 29, 32, 33, 36, 40],
          25: [16, 24, 25, 29, 33, 36, 37, 40, 41]},
    "W": {9: [10, 16, 22, 24, 25, 23], 10: [22, 24, 25, 23], 16: [22, 23],
          24: [26, 27, 30, 31, 34, 35, 37, 38, 39, 41],
          25: [26, 27, 28, 30, 31, 32, 34, 35, 38, 39]},
    "X": {10: [9], 16: [9], 22: [10, 16, 9], 23: [16, 10, 9], 24: [10, 9], 25: [10, 9],
          26: [24, 25], 27: [24, 25], 28: [25], 30: [24, 25], 31: [24, 25], 32: [25],
          34: [24, 25], 35: [24, 25], 37: [24], 38: [24, 25], 39: [24, 25], 41: [24]},
}

TREE_B = {
    "nodeCount": 41,
    "U": {9: [9], 10: [10], 24: [24, 28, 29, 32, 33]},
    "W": {9: [10, 24, 22], 10: [24, 22], 24: [27, 30, 31, 26],
          **{s: _w_b_leaf(s) for s in range(27, 34)}},
    "X": {10: [9], 22: [10, 9], 24: [10, 9], 26: [24], 27: [24], 30: [24], 31: [24],
          34: [27, 28, 29, 30, 31, 32, 33], 35: [28, 29, 30, 31, 32, 33],
          36: [27, 29, 30, 31, 32, 33], 37: [30, 31, 32, 33], 38: [27, 28, 29, 31, 32, 33],
          39: [28, 29, 32, 33], 40: [27, 29, 31, 33]},
}

CASES = {"A": (tree_a, TREE_A), "B": (tree_b, TREE_B)}


def check(lists, count, expect):
    """lists(node, kind) with kind 0..3 = U, V, W, X; compares every node's W
    and X (and the U lists given) with the hand-worked ones."""
    assert count == expect["nodeCount"]
    for n, u in expect["U"].items():
        np.testing.assert_array_equal(lists(n, 0), u, err_msg=f"U({n})")
    for kind, name in ((2, "W"), (3, "X")):
        for n in range(1, count + 1):
            np.testing.assert_array_equal(lists(n, kind), expect[name].get(n, []), err_msg=f"{name}({n})")
