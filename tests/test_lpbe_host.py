"""The linearized Poisson-Boltzmann benchmark problem (problems.lpbe_sphere,
examples/benchmark_LPBE of the reference), pinned without the reference:

* closed forms: the exact pair solves each side's equation (u- constant; u+
  satisfies alpha_2 (kappa^2 u - Laplace u) = 0, by central differences), and
  its jumps across the sphere equal alpha and beta in the reference's sign
  convention (beta = -(mu+ grad u+ - mu- grad u-) . n);
* the discrete equations (oracle records at zoom level 0): the exact solution
  satisfies them to second order away from the interface and to ~1e-3 of the
  residual scale next to it, and planting a sign error in beta or alpha (in the
  oracle's callbacks) raises the residual next to the interface several-fold;
* the grid solve of the same equations (oracle.judge) is closest to the exact
  solution with the correct beta, against beta = 0 and beta of the wrong sign.
"""
import numpy as np

from helpers import grid_points
from jax_dips import problems
from oracle import discretization as D
from oracle import judge as J
from oracle import trainer as T


def _points_on_sphere(m, radius, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(3, m))
    return v / np.linalg.norm(v, axis=0) * radius


def test_exact_solution_closed_forms():
    prob = problems.lpbe_sphere(h=1e-4)
    kappa = prob["kappa"]
    assert abs(kappa - (8.4869e-5 / 80.0) ** 0.5) < 1e-15
    # u- is the constant omega / (4 pi) (1 / (80 (1 + kappa)) - 1 / 2) of a unit sphere
    r_in = _points_on_sphere(16, 0.5)
    assert np.allclose(prob["exact_m_fn"](r_in), 24.0 / (4 * np.pi) * (1 / (80 * (1 + kappa)) - 0.5), rtol=1e-14)
    # u+ solves kappa^2 u - Laplace u = 0 (k+ / mu+ = kappa^2): second differences, step 1e-3
    r_out = _points_on_sphere(16, 1.7, seed=1)
    up = prob["exact_p_fn"]
    e = 1e-3
    lap = sum((up(r_out + e * np.eye(3)[:, i:i + 1]) - 2 * up(r_out) + up(r_out - e * np.eye(3)[:, i:i + 1])) / e ** 2
              for i in range(3))
    assert np.allclose(prob["k_p_fn"](r_out) / prob["mu_p_fn"](r_out), kappa ** 2)
    assert np.max(np.abs(lap - kappa ** 2 * up(r_out))) < 1e-6 * np.max(np.abs(up(r_out)))
    # the boundary value is u+
    assert np.allclose(prob["dirichlet_bc_fn"](r_out), up(r_out))
    # jumps on the sphere: [u] = alpha = g (the Coulomb term of the point charge), and
    # -(mu+ du+/dn - mu- du-/dn) = beta (u- constant; du+/dn by central differences along n)
    r_s = _points_on_sphere(32, 1.0, seed=2)
    n = r_s / np.linalg.norm(r_s, axis=0)
    assert np.allclose(up(r_s) - prob["exact_m_fn"](r_s), prob["alpha_fn"](r_s), rtol=1e-12)
    assert np.allclose(prob["alpha_fn"](r_s), 24.0 / (4 * np.pi * 2.0), rtol=1e-12)
    dup = (up(r_s + e * n) - up(r_s - e * n)) / (2 * e)
    flux_jump = -(80.0 * dup - 2.0 * 0.0)
    assert np.allclose(flux_jump, prob["beta_fn"](r_s), rtol=1e-6)
    assert np.allclose(prob["beta_fn"](r_s), 24.0 / (4 * np.pi), rtol=1e-6)  # omega / (4 pi sigma^2), sigma = 1


def _residual(prob, n):
    axes, P, h, bounds = grid_points(n, -2.5, 2.5)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    rec = T.LevelRecords(P, T.level_spacing(h, 0), bounds, cb, np.float64)
    rows = np.arange(P.shape[0])
    W, nz, X, ph = rec.entries(rows)
    X64 = X.T.astype(np.float64)
    u = np.where(ph, prob["exact_p_fn"](X64), prob["exact_m_fn"](X64))
    r = rec.c[rows].astype(np.float64).copy()
    np.add.at(r, nz[0], W[nz] * u)
    near = np.abs(prob["lvl_set_fn"](P.T.astype(np.float64))) < h[0]
    return r, near, rec


def _planted(h, which):
    prob = problems.lpbe_sphere(h)
    f = prob[which]
    prob[which] = lambda r, f=f: -f(r)
    return prob


def test_exact_solution_satisfies_the_discrete_equations():
    far_max, near_max = [], []
    for n in (24, 32):
        h = 5.0 / (n - 1)
        r, near, _ = _residual(problems.lpbe_sphere(h), n)
        far_max.append(np.abs(r[~near]).max())
        near_max.append(np.abs(r[near]).max())
        rb, _, _ = _residual(_planted(h, "beta_fn"), n)
        ra, _, _ = _residual(_planted(h, "alpha_fn"), n)
        print(n, "far", far_max[-1], "near", near_max[-1], "beta flipped", np.abs(rb[near]).max(), "alpha flipped",
              np.abs(ra[near]).max())
        assert np.abs(rb[near]).max() > 1.7 * near_max[-1]
        assert np.abs(ra[near]).max() > 30 * near_max[-1]
    # second order away from the interface, first order next to it (h ratio 23/31)
    assert far_max[1] < 0.65 * far_max[0] and far_max[1] < 1e-5
    assert near_max[1] < 0.85 * near_max[0] and near_max[1] < 6e-3


def test_grid_solve_is_closest_with_the_correct_beta():
    n = 24
    h = 5.0 / (n - 1)
    axes, P, _, _ = grid_points(n, -2.5, 2.5)
    errs = {}
    for name in ("correct", "zero", "flipped"):
        prob = problems.lpbe_sphere(h)
        if name == "zero":
            prob["beta_fn"] = lambda r: 0.0 * r[0]
        elif name == "flipped":
            prob = _planted(h, "beta_fn")
        _, _, rec = _residual(prob, n)
        u = J.solve(rec, (n, n, n))
        ex = prob["exact_fn"](P.T.astype(np.float64))
        errs[name] = np.linalg.norm(u - ex) / np.linalg.norm(ex)
    print(errs)
    assert errs["correct"] < errs["zero"] < errs["flipped"]
    assert errs["correct"] < 0.03
