"""Shared builders for the parity tests (grids, oracle records, packing)."""
import numpy as np

from jax_dips import problems
from oracle import discretization as D
from oracle import trainer as T


def grid_points(n, lo=-1.0, hi=1.0):
    x, y, z = problems.grid_axes(n, lo, hi)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    h = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]], np.float32)
    bounds = (x.min(), x.max(), y.min(), y.max(), z.min(), z.max())
    return (x, y, z), P, h, bounds


def oracle_level(prob, n, level, dt=np.float64):
    axes, P, h, bounds = grid_points(n)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    rec = T.LevelRecords(P, T.level_spacing(h, level), bounds, cb, dt)
    return axes, P, h, bounds, cb, rec
