"""GPU parity of the hash network model type with the identity encoder (CELU
coordinate MLPs, nn/hash_network.py) against the CPU oracle (oracle/mlp.py,
act="celu"): the tile kernel's loss and gradient, batched training, network
evaluation, and the drop-in trainer with model_dict["model_type"] =
"multiresolution_hash_network" against the oracle epoch loop.

Tolerances as for the tanh networks (test_gpu_tiles.py): float32 against the
float64 oracle, 1e-5 relative on the loss, 1e-4 relative (2-norm) on the
gradient; CELU uses the device expm1f below 0.
"""
import numpy as np
import pytest

from helpers import grid_points, oracle_level
from jax_dips import problems
from oracle import discretization as D
from oracle import mlp as M
from oracle import optim as O
from oracle import trainer as T

LPBE = {"bound": 1.0, "pos_enc": "identity", "N_min": 16, "N_max": 2048, "F": 2, "T": 524288, "tv_scale": 0.0,
        "pos_levels": 16, "layer_widths": [16], "sol_skip_in_layers": [], "sol_act": "linear", "highest_sh_order": 2}
OPT = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


def _glorot(dims, seed):
    from jax_dips.nn.hash_network import HashNetwork

    model = HashNetwork(**dict(LPBE, layer_widths=[dims[1]] * dims[0]))
    return model, model.flatten(model.init(seed))


@pytest.mark.gpu
@pytest.mark.parametrize("dims,level", [((1, 16, 1, 16), 1), ((2, 16, 2, 16), 2)])
def test_celu_tile_loss_grad(dims, level):
    from jax_dips.records import DeviceGrid, DeviceRecords, Engine, pack_dense

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    n = P.shape[0]
    _, theta = _glorot(dims, 5)
    grid, drec = DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, n), rec.h)
    eng = Engine(dims, theta, {}, max_rows=n, max_steps=4, act=1)
    loss, grad = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, rec, np.arange(n), act="celu")
    rel = np.linalg.norm(grad - og) / np.linalg.norm(og)
    print(f"{dims} level {level}: loss gpu {loss:.9e} oracle {ol:.9e} grad rel {rel:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert rel <= 1e-4


@pytest.mark.gpu
def test_celu_eval_matches_oracle():
    from jax_dips.evaluation import evaluate_points

    model, theta = _glorot((1, 16, 1, 16), 9)
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (2000, 3)).astype(np.float32)
    phase = rng.uniform(size=2000) > 0.5
    u, g = evaluate_points(model, theta, X, phase, want_grad=True)
    uo = M.forward(theta.astype(np.float64), model.dims, X.astype(np.float64), phase, act="celu")
    assert np.max(np.abs(u - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))
    eps = 1e-3
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fd = (M.forward(theta, model.dims, X + e, phase, act="celu")
              - M.forward(theta, model.dims, X - e, phase, act="celu")) / (2 * eps)
        assert np.max(np.abs(g[:, a] - fd)) <= 1e-3


@pytest.mark.gpu
def test_trainer_hash_network_matches_oracle(tmp_path):
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    model_dict = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": LPBE}
    model, theta0 = _glorot((1, 16, 1, 16), 42)
    x, y, z = problems.grid_axes(16)
    g = mesh.construct(3)[0](x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model_dict, init_params=theta0)
    st, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, model.dims, theta0, 8, 2048, OPT, act="celu")
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev = tr.model.flatten(tr.params)
    assert set(tr.params["params"]) == {"sol_p", "sol_m"}
    assert np.linalg.norm(dev - th) <= 1e-4 * np.linalg.norm(th)
    phase = cb("phi", P) >= 0
    assert np.max(np.abs(st.solution - M.forward(dev, model.dims, P, phase, act="celu"))) <= 1e-5
