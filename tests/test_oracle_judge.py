"""Pins the oracle's grid-value operator (oracle/judge.py): constant
coefficients, no interface, quadratic manufactured u -> the discrete solution
equals u at the nodes (the 7-point FV stencil is exact for quadratics)."""
import numpy as np

from helpers import grid_points
from oracle import discretization as D
from oracle import judge as J
from oracle import trainer as T


def quad(r):
    return r[0] ** 2 + 2 * r[1] ** 2 - r[2] ** 2 + r[0] * r[1] + 0.3 * r[2]


def test_quadratic_reproduced_exactly():
    mu = 2.5
    c = lambda v: (lambda r: np.zeros_like(r[0]) + v)
    fns = dict(phi=c(-1.0), mu_m=c(mu), mu_p=c(7.0), k_m=c(0.0), k_p=c(0.0),
               f_m=c(-mu * (2 + 4 - 2)), f_p=c(0.0), alpha=c(0.0), beta=c(0.0), g=quad)
    cb = D.Callbacks(fns, np.float64)
    axes, P, h, bounds = grid_points(9)
    rec = T.LevelRecords(P, h, bounds, cb)
    u = J.solve(rec, (9, 9, 9))
    exact = quad(P.astype(np.float64).T)
    assert np.max(np.abs(u - exact)) <= 1e-10
    A, cc = J.operator(rec, (9, 9, 9))
    assert np.max(np.abs(A @ exact + cc)) <= 1e-10
