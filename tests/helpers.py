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


def phase_aware_interp(axes, values, phi_nodes, pts, phase_pts):
    """Interpolate a grid function that jumps across the interface, from nodes of
    the query's own phase only.  values, phi_nodes: (nx, ny, nz); pts (m, 3);
    phase_pts (m,) bool (phi >= 0).  In a cell whose 8 corners share the query's
    phase: trilinear.  In a cut cell: least-squares quadratic (10 terms) through
    the same-phase nodes of the 6x6x6 block around the cell (linear where those
    do not determine a quadratic, nearest same-phase node where they do not
    determine a plane), so the
    interpolant stays second-order up to the interface instead of averaging
    across the jump."""
    x, y, z = (np.asarray(a, np.float64) for a in axes)
    n = np.array([len(x), len(y), len(z)])
    h = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]])
    pts = np.asarray(pts, np.float64)
    pos = np.asarray(phi_nodes >= 0)
    ci = np.stack([np.clip(np.searchsorted(ax, pts[:, a], side="right") - 1, 0, n[a] - 2)
                   for a, ax in enumerate((x, y, z))], axis=1)
    t = np.stack([(pts[:, a] - ax[ci[:, a]]) / (ax[ci[:, a] + 1] - ax[ci[:, a]]) for a, ax in enumerate((x, y, z))],
                 axis=1)
    out = np.empty(len(pts))
    corners = [(a, b, c) for a in (0, 1) for b in (0, 1) for c in (0, 1)]
    for q in range(len(pts)):
        i, j, k = ci[q]
        same = [pos[i + a, j + b, k + c] == phase_pts[q] for a, b, c in corners]
        if all(same):
            v = 0.0
            for a, b, c in corners:
                w = (t[q, 0] if a else 1 - t[q, 0]) * (t[q, 1] if b else 1 - t[q, 1]) * (t[q, 2] if c else 1 - t[q, 2])
                v += w * values[i + a, j + b, k + c]
            out[q] = v
            continue
        rows, vals = [], []
        for a in range(max(i - 2, 0), min(i + 4, n[0])):
            for b in range(max(j - 2, 0), min(j + 4, n[1])):
                for c in range(max(k - 2, 0), min(k + 4, n[2])):
                    if pos[a, b, c] == phase_pts[q]:
                        rows.append(((x[a], y[b], z[c]) - pts[q]) / h)
                        vals.append(values[a, b, c])
        d = np.array(rows).reshape(-1, 3)
        out[q] = vals[int(np.argmin((d ** 2).sum(axis=1)))] if rows else np.nan
        fits = []
        if len(rows) >= 14:
            fits.append(np.column_stack([np.ones(len(d)), d, d ** 2, d[:, 0] * d[:, 1], d[:, 0] * d[:, 2],
                                         d[:, 1] * d[:, 2]]))
        if len(rows) >= 6:
            fits.append(np.column_stack([np.ones(len(d)), d]))
        for A in fits:  # a fit whose nodes do not determine it (all on a few planes) is skipped
            coef, _, rank, _ = np.linalg.lstsq(A, np.array(vals), rcond=None)
            if rank == A.shape[1]:
                out[q] = coef[0]
                break
    return out
