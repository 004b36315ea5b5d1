"""The oracle's learned preconditioner (oracle/precond.py): closed forms and
finite differences.

* s = 0 gives P = 0.5 exactly, so the preconditioned loss is a quarter of the
  plain one and its network gradient a quarter of the plain gradient;
* P lies in (0.5, 0.5 + s);
* the 26 coefficients of an uncut bulk cell are mu*h (faces of its phase), h^3
  and h^2, zero for the other phase;
* the gradient of the preconditioned loss (network and preconditioner
  parameters) against float64 central finite differences.
"""
import numpy as np

from helpers import oracle_level
from jax_dips import problems
from oracle import mlp as M
from oracle import precond as PC
from oracle import trainer as T

DIMS = (1, 1, 2, 10)
WIDTHS = (8, 4)


def _setup(level=1, n=10):
    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, n, level)
    return rec, np.arange(P.shape[0])


def test_scale_zero_is_a_quarter():
    rec, rows = _setup()
    theta = M.init_params(DIMS, seed=1).astype(np.float64)
    phi = PC.init_params(WIDTHS, seed=2).astype(np.float64)
    l0, g0, _ = T.loss_and_grad(theta, DIMS, rec, rows)
    l1, g1, _ = T.loss_and_grad(np.concatenate([theta, phi]), DIMS, rec, rows,
                                precond=dict(widths=WIDTHS, scale=0.0))
    assert np.isclose(l1, 0.25 * l0, rtol=1e-14)
    assert np.allclose(g1[:theta.size], 0.25 * g0, rtol=1e-12, atol=1e-18)
    assert np.all(g1[theta.size:] == 0.0)


def test_range_and_bulk_coefficients():
    rec, rows = _setup()
    phi = PC.init_params(WIDTHS, seed=3).astype(np.float64)
    P, _ = PC.forward(phi, WIDTHS, rec.coeffs, 1.5)
    assert np.all(P > 0.5) and np.all(P < 2.0)
    h = float(rec.h[0])
    C = rec.coeffs
    # an uncut interior point outside the sphere (phase +, mu+ = 10)
    reg = rec.slots["regions"]
    bulk_p = np.nonzero((reg == 1).all(axis=1) & ~rec.boundary)[0][0]
    c = C[bulk_p]
    assert np.allclose(c[1:12:2], 10.0 * h) and np.allclose(c[0:12:2], 0.0)
    assert np.isclose(c[13], h ** 3) and c[12] == 0.0
    assert np.allclose(c[15:26:2], h * h) and np.allclose(c[14:26:2], 0.0)


def test_precond_gradient_matches_finite_differences():
    rec, rows = _setup(level=0, n=8)
    theta = M.init_params(DIMS, seed=4).astype(np.float64)
    phi = PC.init_params(WIDTHS, seed=5).astype(np.float64)
    full = np.concatenate([theta, phi])
    pc = dict(widths=WIDTHS, scale=1.0)
    _, g, _ = T.loss_and_grad(full, DIMS, rec, rows, precond=pc)
    rng = np.random.default_rng(6)
    idx = np.concatenate([rng.choice(theta.size, 12, replace=False),
                          theta.size + rng.choice(phi.size, 20, replace=False)])
    eps = 1e-6
    for i in idx:
        e = np.zeros_like(full)
        e[i] = eps
        lp, _, _ = T.loss_and_grad(full + e, DIMS, rec, rows, precond=pc)
        lm, _, _ = T.loss_and_grad(full - e, DIMS, rec, rows, precond=pc)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd)), (i, fd, g[i])
