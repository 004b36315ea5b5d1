"""Batched training of many problems (dips_batch_train, records.BatchEngine,
solvers.poisson.batch.solve_batch) against single-problem training.

Every problem of a batched run must follow exactly the trajectory that the
single-problem tile path (dips_tile_train_level, itself checked against the
oracle in test_gpu_tiles.py) gives it: the step kernel body, block count and
reduction order are the same, so the per-step losses and final parameters are
compared bit for bit.  One test also runs the batched path against the oracle
loop directly (float32 against float64: 1e-4 relative per step loss, 1e-3 on
the parameters, as in test_gpu_tiles.py).
"""
import numpy as np
import pytest

from helpers import oracle_level
from jax_dips import problems
from oracle import mlp as M
from oracle import optim as O
from oracle import trainer as T

CFG = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


def _level(prob, n, level, batch):
    """Device records of one level built by the geometry kernels."""
    from helpers import grid_points
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid

    axes, P, h, bounds = grid_points(n)
    grid = DeviceGrid(*axes)
    rec = build_level(grid, T.level_spacing(h, level), bounds, problems.oracle_callbacks(prob), batch)
    return grid, rec


CASES = [(problems.sphere(), 12, 1728), (problems.sphere(mu_p=50.0), 16, 1000),
         (problems.star(), 14, 1024), (problems.sphere(radius=0.35, k_m=1.0, k_p=1.0), 20, 4096)]


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(1, 1, 2, 10), (2, 16, 2, 16)])
def test_batch_equals_single_problem_runs(dims):
    from jax_dips.records import BatchEngine, Engine
    import torch

    levels = [_level(p, n, 1, b) for p, n, b in CASES]
    grids = [g for g, _ in levels]
    recs = [r for _, r in levels]
    params = [M.init_params(dims, seed=40 + i) for i in range(len(CASES))]
    epochs = 3
    steps = [epochs * len(r.batch_begin) for r in recs]
    be = BatchEngine(dims, params, CFG, [64] * len(CASES))
    be.train_level(grids, recs, epochs)
    torch.cuda.synchronize()
    for i, (g, r) in enumerate(levels):
        e = Engine(dims, params[i], CFG, max_rows=1, max_steps=64)
        e.train_level(g, r, epochs)
        torch.cuda.synchronize()
        b = be.engines[i]
        print(f"problem {i}: {steps[i]} steps, last loss {float(b.loss_hist[steps[i] - 1]):.6e}")
        assert int(b.count.item()) == int(e.count.item()) == steps[i]
        assert np.array_equal(b.loss_hist[:steps[i]].cpu().numpy(), e.loss_hist[:steps[i]].cpu().numpy()), i
        assert np.array_equal(b.params.cpu().numpy(), e.params.cpu().numpy()), i
        assert np.array_equal(b.m.cpu().numpy(), e.m.cpu().numpy()), i


@pytest.mark.gpu
def test_batch_matches_oracle_steps():
    """Two problems with 2x16 networks, 4 epochs of 4 and 3 batches, against the
    oracle loop with the float32 optax restatement."""
    from jax_dips.records import BatchEngine, DeviceGrid, DeviceRecords, pack_dense

    dims = (2, 16, 2, 16)
    specs = [(problems.sphere(), 12, 432), (problems.star(), 11, 500)]
    grids, recs, orecs, params = [], [], [], []
    for i, (prob, n, B) in enumerate(specs):
        axes, P, h, bounds, cb, rec = oracle_level(prob, n, 0)
        packed = pack_dense(rec.W, rec.c, B)
        grids.append(DeviceGrid(*axes))
        recs.append(DeviceRecords(packed, rec.h))
        orecs.append((rec, P.shape[0], B))
        params.append(M.init_params(dims, seed=60 + i))
    be = BatchEngine(dims, params, CFG, [64, 64])
    be.train_level(grids, recs, 4)
    for i, (rec, n, B) in enumerate(orecs):
        th = params[i].copy()
        st = O.OptState(th.size, np.float32)
        ol = []
        for e in range(4):
            for b0 in range(0, n, B):
                rows = np.arange(b0, min(b0 + B, n))
                l, g, _ = T.loss_and_grad(th.astype(np.float64), dims, rec, rows)
                l, g = l * len(rows) / B, g * len(rows) / B  # the short last batch is divided by B too
                ol.append(l)
                th = O.step(th, g.astype(np.float32), st, CFG, np.float32)
        ol = np.array(ol)
        eng = be.engines[i]
        gl = eng.loss_hist[:len(ol)].cpu().numpy().astype(np.float64)
        gp = eng.params.cpu().numpy()
        print(f"problem {i}: gpu {gl[:3]} oracle {ol[:3]}")
        assert np.all(np.abs(gl - ol) <= 1e-4 * np.abs(ol))
        assert np.linalg.norm(gp - th) <= 1e-3 * np.linalg.norm(th)


@pytest.mark.gpu
def test_solve_batch_matches_trainer_runs(tmp_path):
    """solve_batch over three problems (one with a padded last batch) equals
    three drop-in trainer runs with the same initial parameters."""
    from jax_dips.domain import mesh
    from jax_dips.nn.mlp.MLP import DoubleMLP
    from jax_dips.solvers.poisson import batch, trainer

    model_dict = {"model_type": "mlp", "mlp": {"hidden_layers_m": 2, "hidden_dim_m": 16, "activation_m": "jnp.tanh",
                                               "hidden_layers_p": 2, "hidden_dim_p": 16, "activation_p": "jnp.tanh"}}
    dims = DoubleMLP(**model_dict).dims
    specs = [(problems.sphere(), 12, 1728), (problems.star(), 14, 1000), (problems.sphere(mu_p=100.0), 10, 512)]
    probs, inits = [], []
    for i, (prob, n, B) in enumerate(specs):
        x, y, z = problems.grid_axes(n)
        g = mesh.construct(3)[0](x, y, z)
        d = {k: prob[k] for k in ("lvl_set_fn", "dirichlet_bc_fn", "mu_m_fn", "mu_p_fn", "k_m_fn", "k_p_fn",
                                  "f_m_fn", "f_p_fn", "alpha_fn", "beta_fn")}
        d.update(lvl_gstate=g, tr_gstate=g, eval_gstate=g, batch_size=B)
        probs.append((d, prob, g, B))
        inits.append(M.init_params(dims, seed=80 + i))
    E = 8
    res = batch.solve_batch([d for d, _, _, _ in probs], num_epochs=E, model_dict=model_dict, optimizer_dict=CFG,
                            init_params=inits)
    for i, (d, prob, g, B) in enumerate(probs):
        init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"],
                                prob["mu_m_fn"], prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"],
                                prob["f_p_fn"], prob["alpha_fn"], prob["beta_fn"])
        state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=E, batch_size=B,
                                  checkpoint_dir=str(tmp_path / f"ck{i}"), results_dir=str(tmp_path),
                                  optimizer_dict=CFG, model_dict=model_dict, init_params=inits[i])
        st, es, le = solve_fn(sim_state=state)
        tr = solve_fn.trainer
        print(f"problem {i}: batch {res[i]['loss_epochs'][-3:]} trainer {le[-3:]}")
        assert np.allclose(res[i]["loss_epochs"], le, rtol=1e-6, atol=0)
        assert np.array_equal(tr.model.flatten(res[i]["params"]), tr.model.flatten(tr.params))
        assert np.max(np.abs(res[i]["solution"] - st.solution)) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [(32, 0), (8, 64)])
def test_batch_shape_equals_each_problem_alone(cfg):
    """With a configured shape (fewer blocks per problem and step, a given
    persistent grid) every problem of the batched run equals its run alone
    through the same entry point and shape, bit for bit; the epochs replayed
    from a captured graph equal the launched ones."""
    from jax_dips.records import BatchEngine
    import torch

    dims = (2, 16, 2, 16)
    levels = [_level(p, n, 0, b) for p, n, b in CASES]
    grids = [g for g, _ in levels]
    recs = [r for _, r in levels]
    params = [M.init_params(dims, seed=60 + i) for i in range(len(CASES))]
    epochs = 3
    together = BatchEngine(dims, params, CFG, [64] * len(CASES))
    together.train_level(grids, recs, epochs, cfg=cfg, graph=True)
    launched = BatchEngine(dims, params, CFG, [64] * len(CASES))
    launched.train_level(grids, recs, epochs, cfg=cfg, graph=False)
    torch.cuda.synchronize()
    for i in range(len(CASES)):
        alone = BatchEngine(dims, [params[i]], CFG, [64])
        alone.train_level([grids[i]], [recs[i]], epochs, cfg=cfg, graph=False)
        torch.cuda.synchronize()
        a, b, c = together.engines[i], alone.engines[0], launched.engines[i]
        n = epochs * len(recs[i].batch_begin)
        assert int(a.count.item()) == int(b.count.item()) == int(c.count.item()) == n
        for x, y, z in ((a.params, b.params, c.params), (a.m, b.m, c.m), (a.v, b.v, c.v),
                        (a.loss_hist, b.loss_hist, c.loss_hist)):
            assert torch.equal(x, y) and torch.equal(x, z), i
        print(f"problem {i}: {n} steps, last loss {float(a.loss_hist[n - 1]):.6e}")


@pytest.mark.gpu
def test_batch_autotune_picks_a_measured_shape(tmp_path):
    """The autotuner times every candidate shape from the same state (the
    problems' parameters and step counters are restored after each), caches
    the choice in memory and in its JSON file, and training with the chosen
    shape matches the oracle loop like any other shape."""
    import json

    from jax_dips import records
    from jax_dips.records import BatchEngine
    import torch

    dims = (2, 16, 2, 16)
    levels = [_level(p, n, 1, b) for p, n, b in CASES]
    grids = [g for g, _ in levels]
    recs = [r for _, r in levels]
    params = [M.init_params(dims, seed=70 + i) for i in range(len(CASES))]
    be = BatchEngine(dims, params, CFG, [64] * len(CASES))
    records._BATCH_TUNED.clear()
    path = str(tmp_path / "tune.json")
    cfg = be.autotune(grids, recs, caps=(256, 16), steps=2, cache_path=path)
    assert cfg in {(256, be.tuning["resident_blocks"]), (256, 2 * be.tuning["resident_blocks"]),
                   (16, be.tuning["resident_blocks"]), (16, 2 * be.tuning["resident_blocks"])}
    assert len(be.tuning["candidates_ms"]) == 4
    for e, p in zip(be.engines, params):
        assert int(e.count.item()) == 0 and np.array_equal(e.params.cpu().numpy(), p)
    disk = json.load(open(path))
    assert [tuple(v["cfg"]) for v in disk.values()] == [cfg]
    records._BATCH_TUNED.clear()
    assert be.autotune(grids, recs, caps=(256, 16), cache_path=path) == cfg  # from the file
    be.train_level(grids, recs, 2, cfg=cfg)
    torch.cuda.synchronize()
    for i, (g, r) in enumerate(levels):
        e = be.engines[i]
        assert int(e.count.item()) == 2 * len(r.batch_begin)
        assert np.all(np.isfinite(e.params.cpu().numpy()))
