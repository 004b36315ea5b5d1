"""End-to-end parity of the drop-in trainer (setup -> init_fn -> solve_fn) with
the CPU oracle's epoch loop on the 16^3 sphere workload: same zoom schedule,
batches, optimizer chain and explicit initial parameters.

Tolerance: the device path builds records in FP64 from float64 callback samples
and trains in float32; the oracle trains in float64 arithmetic with float32
parameters and optimizer state.  Per-epoch losses agree to 1e-4 relative over
a short run; the final evaluation matches the oracle network to 1e-5 absolute.
"""
import numpy as np
import pytest

from helpers import grid_points
from jax_dips import problems
from oracle import discretization as D
from oracle import mlp as M
from oracle import trainer as T

MODEL = {"model_type": "mlp", "mlp": {"hidden_layers_m": 1, "hidden_dim_m": 1, "activation_m": "jnp.tanh",
                                      "hidden_layers_p": 2, "hidden_dim_p": 10, "activation_p": "jnp.tanh"}}
OPT = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


def run_trainer(prob, n, epochs, batch, theta0, tmp_path):
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    init_mesh_fn, _ = mesh.construct(3)
    x, y, z = problems.grid_axes(n)
    g = init_mesh_fn(x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=epochs, batch_size=batch,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=MODEL, init_params=theta0)
    new_state, epoch_store, loss_epochs = solve_fn(sim_state=state)
    return g, new_state, epoch_store, loss_epochs, solve_fn.trainer


@pytest.mark.gpu
def test_trainer_matches_oracle_loop(tmp_path):
    prob = problems.sphere()
    dims = (1, 1, 2, 10)
    theta0 = M.init_params(dims, seed=42)
    g, st, es, le, tr = run_trainer(prob, 16, 8, 2048, theta0, tmp_path)
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, dims, theta0, 8, 2048, OPT)
    print("device", le, "oracle", ole)
    assert np.array_equal(es, np.arange(8))
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev_theta = tr.model.flatten(tr.params)
    assert np.linalg.norm(dev_theta - th) <= 1e-4 * np.linalg.norm(th)
    # evaluation on the eval grid = network of the final parameters
    phase = cb("phi", P) >= 0
    u_or = M.forward(dev_theta, dims, P, phase)
    assert np.max(np.abs(st.solution - u_or)) <= 1e-5
    assert st.grad_solution.shape == (P.shape[0], 3) and np.all(np.isfinite(st.grad_normal_solution))


@pytest.mark.gpu
def test_eval_gradients_match_finite_differences():
    from jax_dips.evaluation import evaluate_points
    from jax_dips.nn.mlp.MLP import DoubleMLP

    model = DoubleMLP(**MODEL)
    theta = M.init_params(model.dims, seed=9)
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, (512, 3)).astype(np.float32)
    phase = rng.uniform(size=512) > 0.5
    u, gr = evaluate_points(model, theta, X, phase, want_grad=True)
    assert np.max(np.abs(u - M.forward(theta, model.dims, X, phase))) <= 1e-6
    eps = 1e-3
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fd = (M.forward(theta, model.dims, X + e, phase) - M.forward(theta, model.dims, X - e, phase)) / (2 * eps)
        assert np.max(np.abs(gr[:, a] - fd)) <= 1e-4


@pytest.mark.gpu
def test_trainer_padded_batches(tmp_path):
    """batch_size not dividing N (16^3 = 4096 points, B = 1500): the last batch
    is padded with random points in the data's range on both sides; per-epoch
    losses and parameters follow the oracle loop."""
    prob = problems.sphere()
    dims = (1, 1, 2, 10)
    theta0 = M.init_params(dims, seed=4)
    g, st, es, le, tr = run_trainer(prob, 16, 8, 1500, theta0, tmp_path)
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, dims, theta0, 8, 1500, OPT)
    print("device", le, "oracle", ole)
    assert len(steps) == 8 * 3
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev_theta = tr.model.flatten(tr.params)
    assert np.linalg.norm(dev_theta - th) <= 1e-4 * np.linalg.norm(th)
