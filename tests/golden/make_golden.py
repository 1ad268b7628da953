"""Write the golden fixtures of tests/golden from the CPU oracle.

The reference is MATLAB and cannot run here, and its own tests hold no
vectors; these fixtures freeze the oracle's outputs on small inputs so that
the oracle and the HIP library are both checked against fixed data.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402
from geometry import blobs, cube_sphere  # noqa: E402


def main():
    oracle.build()
    rng = np.random.default_rng(2024)
    # one sphere, NEAREST quadrature: 384 force points, 6144 fine points
    p, fine = cube_sphere(8), cube_sphere(32)
    r, c, w = oracle.nearest_matrix(p, fine)
    NN = sp.csc_matrix((w, (r, c)), shape=(len(fine), len(p)))
    q = rng.standard_normal(p.shape)
    t = oracle.Tree(p, p, nodeCapacity=64, finePoints=fine, NN=NN)
    np.savez_compressed(os.path.join(HERE, "sphere_nearest.npz"), points=p, fine=fine,
                        nn_row=r, nn_col=c, nn_w=w, q=q, eps=0.01, mu=1.0, cap=64, res=6,
                        u=t.matvec(q, 0.01, 1.0, 6, 1), u_reduced=t.reduced_apply(q, 0.01),
                        f_reduced=t.reduced_solve(q, 0.01))
    # adaptive tree with W and X lists, coronaRes 4
    b = blobs(rng)[::2]
    q = rng.standard_normal(b.shape)
    t = oracle.Tree(b, b, nodeCapacity=12)
    np.savez_compressed(os.path.join(HERE, "adaptive_res4.npz"), points=b, q=q, eps=5e-3, mu=1.0,
                        cap=12, res=4, u=t.matvec(q, 5e-3, 1.0, 4, 1))
    # uniform points, distinct targets, coronaRes 6
    s = rng.random((1500, 3))
    tg = np.concatenate([rng.random((700, 3)), s[:100]])
    q = rng.standard_normal(s.shape)
    t = oracle.Tree(s, tg, nodeCapacity=40)
    np.savez_compressed(os.path.join(HERE, "distinct_res6.npz"), points=s, targets=tg, q=q, eps=1e-3,
                        mu=0.8, cap=40, res=6, u=t.matvec(q, 1e-3, 0.8, 6, 1))


if __name__ == "__main__":
    main()
