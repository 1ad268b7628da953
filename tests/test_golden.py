"""The oracle reproduces the committed golden fixtures (CPU)."""
import os

import numpy as np
import scipy.sparse as sp

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name)))


def test_golden_sphere_nearest(oracle):
    g = load("sphere_nearest.npz")
    NN = sp.csc_matrix((g["nn_w"], (g["nn_row"], g["nn_col"])), shape=(len(g["fine"]), len(g["points"])))
    t = oracle.Tree(g["points"], g["points"], nodeCapacity=int(g["cap"]), finePoints=g["fine"], NN=NN)
    u = t.matvec(g["q"], float(g["eps"]), float(g["mu"]), int(g["res"]), 1)
    np.testing.assert_allclose(u, g["u"], rtol=0, atol=1e-13 * np.abs(g["u"]).max())
    # root + 8 leaves: the NEAREST matvec is K(fine -> points) (NN f)
    ud = oracle.direct(g["fine"], NN @ g["q"], g["points"], float(g["eps"]), float(g["mu"]))
    assert np.linalg.norm(g["u"] - ud) / np.linalg.norm(ud) < 1e-13
    np.testing.assert_allclose(t.reduced_apply(g["q"], float(g["eps"])), g["u_reduced"], rtol=0,
                               atol=1e-13 * np.abs(g["u_reduced"]).max())


def test_golden_adaptive_and_distinct(oracle):
    for name in ("adaptive_res4.npz", "distinct_res6.npz"):
        g = load(name)
        tg = g.get("targets", g["points"])
        t = oracle.Tree(g["points"], tg, nodeCapacity=int(g["cap"]))
        u = t.matvec(g["q"], float(g["eps"]), float(g["mu"]), int(g["res"]), 1)
        np.testing.assert_allclose(u, g["u"], rtol=0, atol=1e-12 * np.abs(g["u"]).max(), err_msg=name)
