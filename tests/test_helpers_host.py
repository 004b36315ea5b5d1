"""The phase-aware interpolant used to compare grid solutions off the grid:
exact for functions that are linear on each side of the interface (trilinear
and the least-squares fits both reproduce linear functions), second-order for
a smooth function with a jump."""
import numpy as np

from helpers import phase_aware_interp
from jax_dips import problems


def _setup(n, seed=0):
    x = problems.grid_axes(n)[0].astype(np.float64)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    phi = np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.5
    pts = np.random.default_rng(seed).uniform(-0.95, 0.95, (3000, 3))
    phip = np.linalg.norm(pts, axis=1) - 0.5
    return (x, x, x), (X, Y, Z), phi, pts, phip


def test_piecewise_linear_is_reproduced():
    axes, (X, Y, Z), phi, pts, phip = _setup(12)
    lin_m = lambda a, b, c: 1.0 + 2.0 * a - 0.5 * b + 0.25 * c
    lin_p = lambda a, b, c: -3.0 + 0.7 * a + 1.5 * b - 2.0 * c
    vals = np.where(phi >= 0, lin_p(X, Y, Z), lin_m(X, Y, Z))
    out = phase_aware_interp(axes, vals, phi, pts, phip >= 0)
    want = np.where(phip >= 0, lin_p(*pts.T), lin_m(*pts.T))
    assert np.max(np.abs(out - want)) <= 1e-9


def test_smooth_jump_converges_second_order():
    f_m = lambda a, b, c: np.sin(2 * a) * np.cos(b) + c ** 2
    f_p = lambda a, b, c: np.exp(0.5 * a) - b * c
    errs = []
    for n in (12, 45):
        axes, (X, Y, Z), phi, pts, phip = _setup(n, seed=1)
        vals = np.where(phi >= 0, f_p(X, Y, Z), f_m(X, Y, Z))
        out = phase_aware_interp(axes, vals, phi, pts, phip >= 0)
        errs.append(np.max(np.abs(out - np.where(phip >= 0, f_p(*pts.T), f_m(*pts.T)))))
    print(errs)
    assert errs[1] < errs[0] / 8.0  # h shrinks 3.9x: 15x for second order, 3.9x for first
