"""End-to-end parity of the drop-in trainer (setup -> init_fn -> solve_fn) with
the CPU oracle's epoch loop on the 16^3 sphere workload: same zoom schedule,
batches, optimizer chain and explicit initial parameters.

Tolerance: the device path builds records in FP64 from float64 callback samples
and trains in float32; the oracle trains in float64 arithmetic with float32
parameters and optimizer state.  Per-epoch losses agree to 1e-5 relative over
a short run (1e-4 with a padded last batch); the final evaluation matches the
oracle network to 1e-5 absolute.
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


def mlp_dict(dims):
    Lm, Hm, Lp, Hp = dims
    return {"model_type": "mlp", "mlp": {"hidden_layers_m": Lm, "hidden_dim_m": Hm, "activation_m": "jnp.tanh",
                                         "hidden_layers_p": Lp, "hidden_dim_p": Hp, "activation_p": "jnp.tanh"}}


def run_trainer(prob, n, epochs, batch, theta0, tmp_path, model=MODEL, **kw):
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
                              model_dict=model, init_params=theta0, **kw)
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
    assert np.all(np.abs(le - ole) <= 1e-5 * np.abs(ole))
    dev_theta = tr.model.flatten(tr.params)
    assert np.linalg.norm(dev_theta - th) <= 1e-4 * np.linalg.norm(th)
    # evaluation on the eval grid = network of the final parameters
    phase = cb("phi", P) >= 0
    u_or = M.forward(dev_theta, dims, P, phase)
    assert np.max(np.abs(st.solution - u_or)) <= 1e-5
    assert st.grad_solution.shape == (P.shape[0], 3) and np.all(np.isfinite(st.grad_normal_solution))


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(1, 1, 2, 10), (3, 64, 3, 64), (3, 32, 3, 32), (1, 1, 3, 64)])
def test_eval_gradients_match_finite_differences(dims):
    """u against the float64 oracle network; grad u against central finite
    differences of it (h = 1e-3); d_n u = grad u . n.  The first shape runs the
    per-lane kernel, the others the MFMA tiles (16 points per wavefront, both
    networks where a tile mixes phases; 517 points leave a partial tile)."""
    from jax_dips.evaluation import evaluate_points
    from jax_dips.nn.mlp.MLP import DoubleMLP

    model = DoubleMLP(**mlp_dict(dims))
    theta = M.init_params(model.dims, seed=9)
    rng = np.random.default_rng(0)
    m = 517
    X = rng.uniform(-1, 1, (m, 3)).astype(np.float32)
    phase = rng.uniform(size=m) > 0.5
    phase[:64] = True  # whole tiles of one phase as well as mixed ones
    phase[64:128] = False
    nrm = rng.normal(size=(m, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    u, gr, gn = evaluate_points(model, theta, X, phase, normals=nrm)
    u_or = M.forward(theta, model.dims, X, phase)
    scale = max(1.0, float(np.max(np.abs(u_or))))
    print(dims, "u err", np.max(np.abs(u - u_or)))
    assert np.max(np.abs(u - u_or)) <= 2e-6 * scale
    eps = 1e-3
    fd = np.zeros((m, 3))
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fd[:, a] = (M.forward(theta, model.dims, X + e, phase) - M.forward(theta, model.dims, X - e, phase)) / (2 * eps)
    gscale = max(1.0, float(np.max(np.abs(fd))))
    print(dims, "grad err", np.max(np.abs(gr - fd)))
    assert np.max(np.abs(gr - fd)) <= 1e-4 * gscale
    assert np.max(np.abs(gn - (gr * nrm).sum(axis=1))) <= 1e-6 * gscale


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


@pytest.mark.gpu
def test_trainer_wide_network_matches_oracle_loop(tmp_path):
    """The drop-in trainer with 3x64 networks per phase (MFMA kernels for the
    training steps and the evaluation): per-epoch losses and final parameters
    follow the oracle loop, and solve_fn returns u, grad u and d_n u of the
    trained network on the eval grid."""
    prob = problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0)
    dims = (3, 64, 3, 64)
    theta0 = M.init_params(dims, seed=7)
    g, st, es, le, tr = run_trainer(prob, 16, 4, 2048, theta0, tmp_path, model=mlp_dict(dims))
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, dims, theta0, 4, 2048, OPT)
    print("device", le, "oracle", ole, "rel", np.abs(le - ole) / np.abs(ole))
    assert np.all(np.abs(le - ole) <= 1e-5 * np.abs(ole))
    dev_theta = tr.model.flatten(tr.params)
    assert np.linalg.norm(dev_theta - th) <= 1e-4 * np.linalg.norm(th)
    phase = cb("phi", P) >= 0
    u_or = M.forward(dev_theta, dims, P, phase)
    assert np.max(np.abs(st.solution - u_or)) <= 2e-6 * max(1.0, np.max(np.abs(u_or)))
    assert st.grad_solution.shape == (P.shape[0], 3) and np.all(np.isfinite(st.grad_normal_solution))


@pytest.mark.gpu
def test_restart_after_sigint_matches_uninterrupted_run(tmp_path, monkeypatch):
    """SIGINT sent from the host after epoch 3 of 8 stops training after that
    epoch (the flag of trainer.py:68-78); the checkpoint written at the stop,
    checkpoint_3, resumes with restart=True and the same num_epochs
    (trainer.py:210-220, :327-351).  The resumed run's epoch_store is 3..7, its
    loss_epochs equal the uninterrupted run's epochs 3..7 and its final
    parameters equal the uninterrupted run's, bit for bit (the same kernels
    with fixed-order reductions, the optimizer state and step count restored
    from the checkpoint)."""
    import os
    import signal

    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    dims = (1, 1, 2, 10)
    theta0 = M.init_params(dims, seed=42)
    _, _, es_full, le_full, tr_full = run_trainer(prob, 16, 8, 1024, theta0, tmp_path / "full")
    calls = {"n": 0}
    orig = trainer.Trainer._epoch

    def interrupt_after_third(engine, grid, rec, graph):
        orig(engine, grid, rec, graph)
        calls["n"] += 1
        if calls["n"] == 3:
            os.kill(os.getpid(), signal.SIGINT)

    monkeypatch.setattr(trainer.Trainer, "_epoch", staticmethod(interrupt_after_third))
    part = tmp_path / "part"
    _, _, es1, le1, tr1 = run_trainer(prob, 16, 8, 1024, theta0, part)
    monkeypatch.setattr(trainer.Trainer, "_epoch", staticmethod(orig))
    assert calls["n"] == 3
    assert sorted(os.listdir(part / "ck")) == ["checkpoint_3"]
    assert tr1.steps == 3 * 4
    assert np.array_equal(le1[:3], le_full[:3])
    _, st2, es2, le2, tr2 = run_trainer(prob, 16, 8, 1024, None, tmp_path / "resume", restart=True,
                                        restart_checkpoint_dir=str(part / "ck"))
    print("full", le_full, "resumed", le2)
    assert np.array_equal(es2, np.arange(3, 8))
    assert tr2.steps == 5 * 4
    assert np.array_equal(le2, le_full[3:])
    assert np.array_equal(tr2.model.flatten(tr2.params), tr_full.model.flatten(tr_full.params))
    assert sorted(os.listdir(tmp_path / "resume" / "ck")) == ["checkpoint_8"]


@pytest.mark.gpu
def test_wide_network_eval_at_256_cubed():
    """The 3x64 evaluation (MFMA tiles, forward-mode grad u) on every node of a
    256^3 grid (16.8 M points, the reference's eval grid for its 128^3 runs):
    u at 8,192 sampled points against the float64 oracle network, grad u at 256
    of them against central differences of it."""
    import time

    import torch
    from jax_dips.evaluation import evaluate_points
    from jax_dips.nn.mlp.MLP import DoubleMLP

    model = DoubleMLP(**mlp_dict((3, 64, 3, 64)))
    theta = M.init_params(model.dims, seed=21)
    x = problems.grid_axes(256)[0]
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    del X, Y, Z
    phase = np.linalg.norm(P, axis=1) >= 0.5
    torch.cuda.synchronize()
    t0 = time.time()
    u, g = evaluate_points(model, theta, P, phase, want_grad=True)
    torch.cuda.synchronize()
    print(f"256^3 eval of u and grad u: {time.time() - t0:.2f} s (host copies included)")
    assert u.shape == (P.shape[0],) and g.shape == (P.shape[0], 3)
    rng = np.random.default_rng(1)
    idx = rng.choice(P.shape[0], 8192, replace=False)
    th = theta.astype(np.float64)
    uo = M.forward(th, model.dims, P[idx].astype(np.float64), phase[idx])
    assert np.max(np.abs(u[idx] - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))
    sub = idx[:256]
    h = 1e-3
    for a in range(3):
        e = np.zeros(3)
        e[a] = h
        Q = P[sub].astype(np.float64)
        fd = (M.forward(th, model.dims, Q + e, phase[sub]) - M.forward(th, model.dims, Q - e, phase[sub])) / (2 * h)
        assert np.max(np.abs(g[sub, a] - fd)) <= 1e-3 * max(1.0, np.max(np.abs(fd))), a


@pytest.mark.gpu
def test_trainer_scattered_points_matches_oracle_loop(tmp_path):
    """Scattered training points, as the reference's LPBE example builds its sampled sets
    (examples/benchmark_LPBE/main.py:149-187 with domain/mesh.py:init_gstate_3d_manually): x, y, z
    are the points' coordinates, not axes.  The drop-in trainer places them as a points-only set and
    matches the oracle epoch loop on the same points (losses to 1e-5, parameters to 1e-4 relative), and its
    evaluation at the points matches the oracle network."""
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    dims = (1, 1, 2, 10)
    theta0 = M.init_params(dims, seed=43)
    rng = np.random.default_rng(5)
    P = rng.uniform(-0.95, 0.95, size=(4096, 3)).astype(np.float32)
    d = np.float32(2.0 / 15)
    g_tr = mesh.init_gstate_3d_manually(P[:, 0], P[:, 1], P[:, 2], d, d, d, None, None, None, None, None, None)
    init_mesh_fn, _ = mesh.construct(3)
    g_lvl = init_mesh_fn(*problems.grid_axes(16))
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g_lvl, tr_gstate=g_tr, eval_gstate=g_tr, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=MODEL, init_params=theta0)
    st, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    h = np.array([d, d, d], np.float32)
    bounds = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, _, _ = T.train(P, h, bounds, cb, dims, theta0, 8, 2048, OPT)
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-5 * np.abs(ole))
    dev = tr.model.flatten(tr.params)
    assert np.linalg.norm(dev - th) <= 1e-4 * np.linalg.norm(th)
    phase = cb("phi", P) >= 0
    assert np.max(np.abs(np.asarray(st.solution) - M.forward(dev, dims, P, phase))) <= 1e-5
