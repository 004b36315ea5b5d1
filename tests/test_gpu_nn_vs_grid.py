"""What the trained network adds over the method without it, measured like for like.

The same cut-cell discretisation at the n^3 training nodes is (a) solved for grid
values by BiCGSTAB (GridJudge, spacing h) and (b) used to train the two-phase
network through the zoom schedule (trainer.py:502-591 with
data_management.py:319-326: the last three quarters of the epochs train on the
h/2, h/4 and h/8 stencils of the same nodes).  Both are compared with the exact
solution
* at the n^3 training nodes, where the grid solve needs no interpolation, and
* at 20,000 uniform off-grid points, where the grid solve is interpolated from
  nodes of the query's own phase (helpers.phase_aware_interp: trilinear in
  uncut cells, least-squares quadratic in cut cells; pinned in
  test_helpers_host.py), so the interpolant does not smear the jump.
The grid solve on the 2n-1 grid (spacing h/2) is printed for scale.

Measured on an MI355X at n = 24, 3x64/3x64 tanh networks, full batch
(L-inf / L2, network then grid solve):

  sphere (mu 1 / 10), 40,000 epochs, lr 3e-3:  nodes 0.046 / 0.0072 vs 0.136 / 0.0096;
      off-grid 0.045 / 0.0082 vs 0.127 / 0.0097  (h/2 grid solve at the nodes 0.064 / 0.0048)
  variable-mu sphere (experiment_configs.py:21-186), 40,000 epochs, lr 1e-3:
      nodes 0.060 / 0.0051 vs 0.189 / 0.0114; off-grid 0.059 / 0.0059 vs 0.155 / 0.0117
      (h/2 grid solve 0.123 / 0.0053)
  star (r0 0.5, amplitude 0.1, 5 x 3 lobes, mu 1 / 100), 20,000 epochs, lr 1e-3:
      nodes 0.18 / 0.016 vs 1.83 / 0.115; off-grid 0.20 / 0.019 vs 1.52 / 0.115
      (h/2 grid solve 0.89 / 0.060)

The gain comes from the finer stencils of the later zoom levels: the network is
trained on h/8 stencils at the n^3 nodes, which no n^3 grid solve can use.  The
budget matters: on the star, 40,000 and 80,000 epochs at lr 1e-3 end at L-inf
0.40 and 0.72 at the nodes (the loss still falls), so the epoch count is part of
each case.  The 1x1/2x10 default network at 2,000 epochs is worse than the grid
solve on the sphere (nodes 0.25 / 0.033).
"""
import numpy as np
import pytest

from helpers import phase_aware_interp
from jax_dips import problems

CASES = [("sphere", 40000, 3e-3), ("ref_sphere", 40000, 1e-3), ("star", 20000, 1e-3)]


def _errs(e):
    return float(np.abs(e).max()), float(np.sqrt(np.mean(e ** 2)))


@pytest.mark.gpu
@pytest.mark.parametrize("name,epochs,lr", CASES)
def test_network_against_grid_solve(tmp_path, name, epochs, lr):
    from jax_dips.domain import mesh
    from jax_dips.judge import GridJudge
    from jax_dips.solvers.poisson import trainer

    prob = getattr(problems, name)()
    n = 24
    axes = problems.grid_axes(n)
    g = mesh.construct(3)[0](*axes)
    R = np.asarray(g.R, np.float64)
    exact = prob["exact_fn"](R.T)
    cbs = problems.oracle_callbacks(prob)
    ug, _, rr = GridJudge(g, g, cbs).solve(rtol=1e-7)
    assert rr <= 1e-6
    ug = np.asarray(ug, np.float64)
    grid_nodes = _errs(ug - exact)
    x2 = problems.grid_axes(2 * n - 1)[0]
    g2 = mesh.construct(3)[0](x2, x2, x2)
    ug2, _, _ = GridJudge(g2, g2, cbs).solve(rtol=1e-7)
    fine_nodes = _errs(np.asarray(ug2, np.float64).reshape((2 * n - 1,) * 3)[::2, ::2, ::2].ravel() - exact)

    h = float(axes[0][1] - axes[0][0])
    pts = np.random.default_rng(0).uniform(-1 + 0.5 * h, 1 - 0.5 * h, (20000, 3))
    phase = np.asarray(prob["lvl_set_fn"](pts.T)) >= 0
    exact_pts = prob["exact_fn"](pts.T)
    phi_nodes = np.asarray(prob["lvl_set_fn"](R.T), np.float64).reshape(n, n, n)
    grid_off = _errs(phase_aware_interp(axes, ug.reshape(n, n, n), phi_nodes, pts, phase) - exact_pts)

    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    model = {"model_type": "mlp", "mlp": {"hidden_layers_m": 3, "hidden_dim_m": 64, "activation_m": "jnp.tanh",
                                          "hidden_layers_p": 3, "hidden_dim_p": 64, "activation_p": "jnp.tanh"}}
    opt = {"optimizer_name": "custom", "learning_rate": lr,
           "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    st, solve = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=epochs, batch_size=n ** 3,
                        checkpoint_dir=str(tmp_path), results_dir=str(tmp_path), optimizer_dict=opt,
                        model_dict=model)
    st2, _, loss = solve(st)
    nn_nodes = _errs(np.asarray(st2.solution, np.float64) - exact)
    u_off = solve.trainer.evaluate_solution_fn(solve.trainer.params, pts)
    nn_off = _errs(np.asarray(u_off, np.float64) - exact_pts)
    print(f"{name}: network nodes {nn_nodes} off-grid {nn_off}; grid solve nodes {grid_nodes} "
          f"off-grid (phase-aware) {grid_off}; h/2 grid solve nodes {fine_nodes}; loss ratio {loss[-1] / loss[0]:.2e}")
    assert loss[-1] < 1e-4 * loss[0]
    # L-inf: the network at least 1.6x better than the grid solve, at the nodes and off the grid
    assert nn_nodes[0] < 0.6 * grid_nodes[0] and nn_off[0] < 0.6 * grid_off[0]
    # L2: better as well (0.75x on the sphere, the closest case)
    assert nn_nodes[1] < 0.9 * grid_nodes[1] and nn_off[1] < 0.95 * grid_off[1]


@pytest.mark.gpu
def test_network_on_the_lpbe_sphere(tmp_path):
    """The reference's LPBE benchmark (problems.lpbe_sphere, known answer) at 24^3 on
    [-2.5, 2.5]^3: the drop-in trainer with 3x64 networks reaches the accuracy of the
    exact solve of the same discrete equations (oracle.judge, sparse direct), within
    a factor of 2 in L-inf and L2 at the nodes; the exact solve itself is ~2% off
    the analytic solution (the grid-value form of the jump conditions, see
    test_lpbe_host.py)."""
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer
    from oracle import discretization as D
    from oracle import judge as J
    from oracle import trainer as T

    n = 24
    axes = problems.grid_axes(n, -2.5, 2.5)
    h = float(axes[0][1] - axes[0][0])
    prob = problems.lpbe_sphere(h)
    g = mesh.construct(3)[0](*axes)
    R = np.asarray(g.R, np.float64)
    exact = prob["exact_fn"](R.T)
    P = R.astype(np.float32)
    hh = np.array([h, h, h], np.float32)
    bounds = (-2.5, 2.5, -2.5, 2.5, -2.5, 2.5)
    rec = T.LevelRecords(P, T.level_spacing(hh, 0), bounds,
                         D.Callbacks(problems.oracle_callbacks(prob), np.float64), np.float64)
    grid_nodes = _errs(J.solve(rec, (n, n, n)) - exact)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    model = {"model_type": "mlp", "mlp": {"hidden_layers_m": 3, "hidden_dim_m": 64, "activation_m": "jnp.tanh",
                                          "hidden_layers_p": 3, "hidden_dim_p": 64, "activation_p": "jnp.tanh"}}
    opt = {"optimizer_name": "custom", "learning_rate": 1e-3,
           "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    st, solve = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=4000, batch_size=n ** 3,
                        checkpoint_dir=str(tmp_path), results_dir=str(tmp_path), optimizer_dict=opt,
                        model_dict=model)
    st2, _, loss = solve(st)
    nn_nodes = _errs(np.asarray(st2.solution, np.float64) - exact)
    print(f"lpbe: network nodes {nn_nodes}; exact solve of the discrete equations {grid_nodes}; "
          f"loss ratio {loss[-1] / loss[0]:.2e}")
    assert loss[-1] < 1e-3 * loss[0]
    assert nn_nodes[0] < 2.0 * grid_nodes[0] and nn_nodes[1] < 2.0 * grid_nodes[1]
