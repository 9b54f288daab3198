"""Writes the committed fixtures of tests/golden/ (run in the build container; numpy + scipy).

poisson_block.npz: an FP64 sparse direct solve (scipy.sparse.linalg.spsolve) of the Dirichlet-0
7-point Poisson problem L p = div on a small irregular leaf-dense domain, assembled here from the
coordinates alone, plus the iteration count scipy's own CG needs to 1e-8. It pins the oracle's and
the GPU's CG/RBGS solves without either of them being involved in making the fixture.
"""
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))


def canonical(origins):
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from hnanosolver_amd.topology import coords_from_leaves, leaf_order
    return coords_from_leaves(origins[leaf_order(origins)])


def poisson(c, dx):
    lut = {tuple(p): i for i, p in enumerate(c)}
    rows, cols, vals = [], [], []
    for i, p in enumerate(c):
        rows.append(i); cols.append(i); vals.append(6.0)
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            j = lut.get((p[0] + d[0], p[1] + d[1], p[2] + d[2]))
            if j is not None:
                rows.append(i); cols.append(j); vals.append(-1.0)
    return sp.csr_matrix((vals, (rows, cols)), shape=(len(c), len(c))) / (dx * dx)


def main():
    org = np.asarray([[0, 0, 0], [8, 0, 0], [16, 0, 0], [0, 8, 0], [8, 8, 0], [0, 0, 8], [8, 0, 8], [-8, 0, 0],
                      [-8, -8, 0], [16, 8, 8]], np.int32)
    c = canonical(org)
    dx = float(np.float32(0.1))
    x = c.astype(np.float64)
    div = (np.sin(0.4 * x[:, 0]) * np.cos(0.3 * x[:, 1]) + 0.2 * np.cos(0.5 * x[:, 2])).astype(np.float32)
    A = poisson(c, dx)                      # A = -L
    b = -div.astype(np.float64)
    p = spla.spsolve(A.tocsc(), b)
    its = []
    spla.cg(A, b, rtol=1e-8, maxiter=10000, callback=lambda xk: its.append(1))
    np.savez_compressed(os.path.join(HERE, "poisson_block.npz"), coords=c, div=div, p=p, dx=dx,
                        scipy_cg_iters_1e8=len(its))
    print("poisson_block.npz: N=%d, scipy CG iterations to 1e-8: %d" % (len(c), len(its)))


if __name__ == "__main__":
    main()
