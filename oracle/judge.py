"""Grid-value operator and its direct solve — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates compute_Ax_and_b_discrete_fn / residual_vector (jax_dips/solvers/
poisson/discretization.py:522-741, trainer.py:883-890) with grid values read at
the nodes and, beyond the box, through the linear ghost layer of
add_ghost_layer_3d (domain/interpolate.py:762-816): u[-1] = 2 u[0] - u[1]
per axis.  The residual rows are the zoom-0 records of oracle.trainer with the
network replaced by node values (the phase of an entry plays no role).

Pinned by tests/test_oracle_judge.py: for constant coefficients without an
interface the operator is the textbook 7-point stencil, exact for quadratic u,
so the direct solve reproduces a quadratic manufactured solution to round-off.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import discretization as D


def _ghost_terms(idx, n):
    if idx < 0:
        return [(0, 2.0), (1, -1.0)]
    if idx >= n:
        return [(n - 1, 2.0), (n - 2, -1.0)]
    return [(idx, 1.0)]


def operator(rec, shape):
    """(A, c): residual r = A u + c as a scipy CSR matrix over grid nodes."""
    nx, ny, nz = shape
    n = nx * ny * nz
    rows, cols, vals = [], [], []
    W = rec.W
    for pt in range(n):
        i, rem = divmod(pt, ny * nz)
        j, k = divmod(rem, nz)
        for key in np.nonzero(W[pt])[0]:
            o = D.OFFS125[key // 2]
            for ii, wi in _ghost_terms(i + o[0], nx):
                for jj, wj in _ghost_terms(j + o[1], ny):
                    for kk, wk in _ghost_terms(k + o[2], nz):
                        rows.append(pt)
                        cols.append((ii * ny + jj) * nz + kk)
                        vals.append(W[pt, key] * wi * wj * wk)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    return A, np.asarray(rec.c, np.float64)


def solve(rec, shape):
    A, c = operator(rec, shape)
    return spla.spsolve(A.tocsc(), -c)
