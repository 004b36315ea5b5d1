"""Pins the oracle's residual records (oracle/discretization.py, oracle/trainer.py).

Closed forms and identities, each chosen so that a dropped term, a wrong sign
or a transposed index fails:
* uncut interior point, constant mu and k: the textbook 7-point FV stencil
  r = u_c - mu h sum(u_nb)/(6 mu h + k h^3) - f h^3/(6 mu h + k h^3);
* box points: r = u(p) - g(p);
* all 27 regression weights active: D = X^T / (18 h^2); a rank-deficient mask
  keeps the pinv cutoff semantics (zero direction dropped);
* linear u is differentiated exactly by D for any full-rank mask;
* mu- = mu+: zeta = gamma = 0, the ghost value is u_c -+ (alpha + delta beta/mu);
* the reference's per-step pattern (every stencil slot evaluated separately,
  oracle/refpattern.py) gives the same loss as the merged records;
* the loss gradient matches central finite differences of the loss.
"""
import numpy as np
import pytest

from jax_dips import problems
from oracle import discretization as D
from oracle import mlp as M
from oracle import refpattern
from oracle import trainer as T
from helpers import grid_points


def const_problem(phi_val, mu_m=1.7, mu_p=2.3, k_m=0.4, k_p=0.9, f_m=1.3, f_p=-0.6):
    c = lambda v: (lambda r: np.zeros_like(r[0]) + v)
    return dict(phi=c(phi_val), mu_m=c(mu_m), mu_p=c(mu_p), k_m=c(k_m), k_p=c(k_p), f_m=c(f_m), f_p=c(f_p),
                alpha=c(0.0), beta=c(0.0), g=lambda r: r[0] + 2 * r[1] - r[2])


@pytest.mark.parametrize("side", [-1, 1])
def test_uncut_stencil_closed_form(side):
    axes, P, h, bounds = grid_points(9)
    mu, k, f = (1.7, 0.4, 1.3) if side < 0 else (2.3, 0.9, -0.6)
    cb = D.Callbacks(const_problem(side * 0.5), np.float64)
    rec = T.LevelRecords(P, h, bounds, cb)
    hh = float(h[0])
    diag = 6 * mu * hh + k * hh ** 3
    inner = ~rec.boundary
    W = rec.W[inner]
    ph = 1 if side > 0 else 0
    centre = 2 * D.offset_code(np.array([0, 0, 0])) + ph
    assert np.allclose(W[:, centre], 1.0, rtol=1e-12)
    for e in D.STENCIL[1:]:
        key = 2 * D.offset_code(e) + ph
        assert np.allclose(W[:, key], -mu * hh / diag, rtol=1e-10)
    assert np.allclose((W != 0).sum(axis=1), 7)
    assert np.allclose(rec.c[inner], -f * hh ** 3 / diag, rtol=1e-10)
    # box points: u - g
    g = P[rec.boundary, 0] + 2 * P[rec.boundary, 1] - P[rec.boundary, 2]
    assert np.allclose(rec.c[rec.boundary], -g, atol=1e-6)
    assert np.all((rec.W[rec.boundary] != 0).sum(axis=1) == 1)


def test_regression_full_mask_and_rank_deficient():
    h = np.array([0.1, 0.1, 0.1], np.float32)
    X = (D.CUBE * h).astype(np.float64)
    Dm, Dp = D.regression_rows(np.ones((1, 27)), h, np.float64)
    assert np.allclose(Dp[0], X.T / (18 * float(h[0]) ** 2), rtol=1e-6)
    assert np.all(Dm == 0)
    # only the z = 0 layer positive: M is singular in z, pinv drops that direction
    phi = np.where(D.CUBE[:, 2] == 0, 1.0, -1.0)[None]
    Dm, Dp = D.regression_rows(phi, h, np.float64)
    assert np.allclose(Dp[0, 2], 0.0, atol=1e-9)
    lay = D.CUBE[:, 2] == 0
    Xl = X[lay][:, :2]
    ref = np.linalg.solve(Xl.T @ Xl, Xl.T)
    assert np.allclose(Dp[0, :2][:, lay], ref, rtol=1e-6)
    # linear u reproduced for a random full-rank mask
    rng = np.random.default_rng(0)
    phi = rng.normal(size=(20, 27))
    phi[:, 13] = 1.0
    Dm, Dp = D.regression_rows(phi, h, np.float64)
    grad = np.array([0.3, -1.2, 0.7])
    u = X @ grad
    for i in range(20):
        for Dx, mask in ((Dp[i], phi[i] > 0), (Dm[i], phi[i] < 0)):
            Xm = X[mask]
            if np.linalg.matrix_rank(Xm) == 3:
                assert np.allclose(Dx @ u, grad, rtol=1e-6)


def test_equal_mu_ghost_value():
    """mu- = mu+ -> zeta = gamma = 0: u- (delta > 0) = u_c - (alpha + delta beta/mu)(r_G)."""
    prob = problems.sphere(radius=0.5, mu_m=3.0, mu_p=3.0)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    h = np.array([0.1, 0.1, 0.1], np.float32)
    rng = np.random.default_rng(1)
    s = rng.normal(size=(400, 3))
    s = (s / np.linalg.norm(s, axis=1, keepdims=True) * rng.uniform(0.47, 0.53, size=(400, 1))).astype(np.float32)
    a_m, a_p, b_m, b_p, aux = D.u_mp_forms(s, h, cb, np.float64)
    cut = aux["cut"]
    assert cut.size > 100
    off13 = np.ones(27, bool)
    off13[13] = False
    assert np.allclose(a_m[cut][:, off13], 0, atol=1e-12) and np.allclose(a_p[cut][:, off13], 0, atol=1e-12)
    assert np.allclose(a_m[cut][:, 13], 1) and np.allclose(a_p[cut][:, 13], 1)
    d = aux["delta"][cut]
    rG = aux["r_proj"]
    J = cb("alpha", rG) + d * cb("beta", rG) / 3.0
    pos = d > 0
    assert np.allclose(b_m[pos], -J[pos]) and np.allclose(b_p[pos], 0)
    assert np.allclose(b_p[~pos], J[~pos]) and np.allclose(b_m[~pos], 0)


def test_reference_pattern_equals_records_and_gradient_fd():
    prob = problems.sphere(radius=0.45, mu_m=1.0, mu_p=4.0)
    axes, P, h, bounds = grid_points(10)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    dims = (1, 3, 2, 6)
    theta = M.init_params(dims, seed=2).astype(np.float64)
    rec = T.LevelRecords(P, h, bounds, cb)
    rows = np.arange(P.shape[0])
    loss, grad, _ = T.loss_and_grad(theta, dims, rec, rows)
    loss_ref, grad_ref = refpattern.step(theta, dims, P, h, bounds, cb, dt=np.float64)
    assert abs(loss - loss_ref) <= 1e-6 * loss
    assert np.linalg.norm(grad - grad_ref) <= 1e-5 * np.linalg.norm(grad)
    rng = np.random.default_rng(3)
    for i in rng.choice(theta.size, 12, replace=False):
        e = np.zeros_like(theta)
        e[i] = 1e-3
        lp, _, _ = T.loss_and_grad(theta + e, dims, rec, rows)
        lm, _, _ = T.loss_and_grad(theta - e, dims, rec, rows)
        fd = (lp - lm) / 2e-3
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])) + 1e-3 * abs(grad[i]), (i, fd, grad[i])


def test_reference_pattern_with_hashgrid_equals_records():
    """The per-step pattern with the hashgrid network (the hashgrid workload's
    CPU baseline) gives the loss and gradient of the merged records."""
    from oracle import hashgrid as HG

    prob = problems.sphere(radius=0.45, mu_m=1.0, mu_p=4.0)
    axes, P, h, bounds = grid_points(8)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    grid = HG.HashGrid(4, 2 ** 10, 2, 4, 32, 1.0)
    theta = HG.init_params(grid, [16], seed=2).astype(np.float64)
    theta[2 * HG.mlp_size(8, [16]):] *= 1e4
    rec = T.LevelRecords(P, h, bounds, cb)
    loss, grad, _ = HG.loss_and_grad(theta, grid, [16], rec, np.arange(P.shape[0]))
    loss_ref, grad_ref = refpattern.step(theta, None, P, h, bounds, cb, dt=np.float64, net=HG.Network(grid, [16]))
    assert abs(loss - loss_ref) <= 1e-6 * loss
    assert np.linalg.norm(grad - grad_ref) <= 1e-5 * np.linalg.norm(grad)


def test_beta_not_finite_at_origin_zeroes_crossed_rhs():
    """Padding triangles sit at the origin: a beta that is NaN there makes the
    interface integral NaN and nan_to_num(rhs/diag) = 0 on every crossed cell."""
    prob = problems.sphere()
    fns = problems.oracle_callbacks(prob)
    good = D.Callbacks(fns, np.float64)
    bad_fns = dict(fns)
    bad_fns["beta"] = lambda r: np.where((r[0] == 0) & (r[1] == 0) & (r[2] == 0), np.nan, fns["beta"](r))
    bad = D.Callbacks(bad_fns, np.float64)
    axes, P, h, bounds = grid_points(12)
    s1 = D.point_slots(P, h, bounds, good)
    s2 = D.point_slots(P, h, bounds, bad)
    cut = (s1["geo"]["region"] == 0) & ~s1["boundary"]
    assert cut.sum() > 10
    lhs_part = s1["c"] + np.nan_to_num(s1["rhs"] / s1["diag"])
    assert np.allclose(s2["c"][cut], lhs_part[cut])
    assert np.allclose(s2["c"][~cut], s1["c"][~cut])


def _linear(c0, g):
    g = np.asarray(g, np.float64)
    return lambda r: c0 + g[0] * r[0] + g[1] * r[1] + g[2] * r[2]


@pytest.mark.parametrize("mbig,pos", [(False, True), (False, False), (True, True), (True, False)])
def test_unequal_mu_ghost_values_closed_form(mbig, pos):
    """The four ghost-value branches of get_u_mp_by_regression_at_point_fn
    (discretization.py:453-514) for mu- != mu+, pinned by closed forms.

    Planar phi = n.(x - x0), two arbitrary linear functions u-, u+ (the network
    values U_q are u+ where phi >= 0, u- elsewhere), linear alpha, beta, mu-,
    mu+.  Least squares reproduces a linear field exactly, so with
    D(U - u_c) = grad u on the matching side and c = n . D 1 (the fitted
    gradient of the constant 1, computed here with numpy.linalg.lstsq on the
    masked stencil rows, not with the oracle's pinv), Delta = u+(s) - u-(s),
    J- = alpha + delta beta/mu- and J+ = alpha + delta beta/mu+ at
    r_G = s - delta n:
      mu- <= mu+, delta > 0:  u- = u+(s) - k dn u+ - J-,       k = (mu+ - mu-)/mu- delta
      mu- <= mu+, delta <= 0: u+ = u-(s) + (z Delta - k dn u+ + J-)/(1 + z),  z = -k c+
      mu- > mu+,  delta > 0:  u- = u+(s) - (k dn u- + z Delta)/(1 - z) - (1 - 2z)/(1 - z) J+,
                              k = (mu+ - mu-)/mu+ delta, z = -k c-
      mu- > mu+,  delta <= 0: u+ = u-(s) - k dn u- + J+
    and the other value is u_c.  The first-order term has the opposite sign to
    a Taylor expansion from r_G (u-(s) = u+(s) + k dn u+ - J- when the jump
    conditions hold), which the oracle reproduces as the reference writes it."""
    rng = np.random.default_rng(11 + 2 * mbig + pos)
    ntest = 0
    for trial in range(60):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        x0 = rng.uniform(-0.3, 0.3, 3)
        hs = np.float32(rng.uniform(0.03, 0.12))
        h = np.array([hs, hs, hs], np.float32)
        delta_want = rng.uniform(0.05, 0.45) * float(hs) * (1 if pos else -1)
        tang = rng.normal(size=3)
        tang -= (tang @ nrm) * nrm
        s = (x0 + delta_want * nrm + 0.2 * tang).astype(np.float32)[None]
        gm, gp = rng.normal(size=3), rng.normal(size=3)
        cm0, cp0 = rng.normal(), rng.normal()
        # mu linear and positive near s, ordered as the branch needs at s
        lo, hi = rng.uniform(1.0, 2.0), rng.uniform(3.0, 8.0)
        mu_m0, mu_p0 = (hi, lo) if mbig else (lo, hi)
        gmu_m, gmu_p = rng.normal(size=3) * 0.3, rng.normal(size=3) * 0.3
        sv = s[0].astype(np.float64)
        fns = dict(phi=lambda r: nrm[0] * (r[0] - x0[0]) + nrm[1] * (r[1] - x0[1]) + nrm[2] * (r[2] - x0[2]),
                   mu_m=_linear(mu_m0 - gmu_m @ sv, gmu_m), mu_p=_linear(mu_p0 - gmu_p @ sv, gmu_p),
                   alpha=_linear(rng.normal(), rng.normal(size=3)), beta=_linear(rng.normal(), rng.normal(size=3)))
        cb = D.Callbacks(fns, np.float64)
        a_m, a_p, b_m, b_p, aux = D.u_mp_forms(s, h, cb, np.float64)
        if aux["cut"].size != 1:
            continue
        pos_q = (s[0][None, :] + (D.CUBE * h[None, :]).astype(np.float32)).astype(np.float32).astype(np.float64)
        phi_q = cb("phi", pos_q)
        um_f, up_f = _linear(cm0, gm), _linear(cp0, gp)
        U = np.where(phi_q >= 0, up_f(pos_q.T), um_f(pos_q.T))
        u_m, u_p = a_m[0] @ U + b_m[0], a_p[0] @ U + b_p[0]
        delta = float(cb("phi", s)[0])
        assert (delta > 0) == pos
        X = (D.CUBE * h[None, :]).astype(np.float32).astype(np.float64)
        side = phi_q > 0 if not mbig else phi_q < 0
        if np.linalg.matrix_rank(X[side]) < 3:
            continue
        c1 = float(nrm @ np.linalg.lstsq(X[side], np.ones(int(side.sum())), rcond=None)[0])
        rG = (sv - delta * nrm)[:, None]
        al, be = fns["alpha"](rG)[0], fns["beta"](rG)[0]
        Jm = al + delta * be / fns["mu_m"](rG)[0]
        Jp = al + delta * be / fns["mu_p"](rG)[0]
        mm, mp = fns["mu_m"](sv[:, None])[0], fns["mu_p"](sv[:, None])[0]
        assert (mm > mp) == mbig
        ups, ums = up_f(sv[:, None])[0], um_f(sv[:, None])[0]
        Dl = ups - ums
        if not mbig:
            k = (mp - mm) / mm * delta
            z = -k * c1
            if pos:
                want_m, want_p = ups - k * (nrm @ gp) - Jm, ups
            else:
                want_m, want_p = ums, ums + (z * Dl - k * (nrm @ gp) + Jm) / (1 + z)
        else:
            k = (mp - mm) / mp * delta
            z = -k * c1
            if pos:
                want_m, want_p = ups - (k * (nrm @ gm) + z * Dl) / (1 - z) - (1 - 2 * z) / (1 - z) * Jp, ups
            else:
                want_m, want_p = ums, ums - k * (nrm @ gm) + Jp
        scale = 1.0 + abs(want_m) + abs(want_p)
        assert abs(u_m - want_m) <= 2e-5 * scale, (trial, u_m, want_m)
        assert abs(u_p - want_p) <= 2e-5 * scale, (trial, u_p, want_p)
        ntest += 1
    assert ntest >= 20
