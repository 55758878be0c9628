"""Writes the golden fixtures of tests/golden from the CPU oracle.

bmat.npz     reference gradients of class 0 for trig p=3 (12-point rule) and
             tet p=4 (125-point rule): [Q][nd][dim]
trig64_p3.npz  CSR pattern of the 64x64-square, 2-trigs-per-square, p=3 space and y = A x for
             x ~ U[-1,1] from numpy's default_rng(0)

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import fe_oracle as O  # noqa: E402


def main():
    out = {}
    for key, (mesh, p, order) in {"trig_p3_q12": (O.structured_2d(2), 3, 6),
                                  "tet_p4_q125": (O.structured_3d(2), 4, 8)}.items():
        pr = O.Problem(mesh, p, intorder=order)
        cls = O.class_number(mesh["dim"], mesh["els"])
        e = int(np.where(cls == 0)[0][0])
        _, g = pr.shapes()
        out[key] = g[e]
    np.savez_compressed(os.path.join(HERE, "bmat.npz"), **out)

    pr = O.Problem(O.structured_2d(64), 3, intorder=6)
    rp, c, v = pr.assemble()
    x = np.random.default_rng(0).uniform(-1, 1, pr.ndof)
    import scipy.sparse as sp
    y = sp.csr_matrix((v, c, rp), shape=(pr.ndof, pr.ndof)) @ x
    np.savez_compressed(os.path.join(HERE, "trig64_p3.npz"), rowptr=rp, cols=c, y=y)


if __name__ == "__main__":
    main()
