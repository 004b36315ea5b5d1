"""Epochs replayed from a captured HIP graph equal the same epochs launched
from the host, bit for bit: loss history, parameters, Adam moments and the
step counter after every replay.  Between replays the program does its own
eager work and synchronizes (reads the state back), as a training loop does.
The captured launches are the tile step + optimizer kernels (1x1/2x10) and
the five MFMA step kernels (3x64, entries grouped by position); no memset or
copy is recorded.  The drop-in trainer with and without graphs gives the same
loss curve and parameters."""
import numpy as np
import pytest

from helpers import grid_points
from jax_dips import problems

OPT = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


def state(eng):
    import torch

    torch.cuda.synchronize()
    return [t.clone() for t in (eng.params, eng.m, eng.v, eng.count, eng.loss_hist)]


@pytest.mark.gpu
@pytest.mark.parametrize("dims,level,batch", [((1, 1, 2, 10), 1, 1024), ((3, 64, 3, 64), 0, 2048)])
def test_graph_replay_equals_eager_epochs(dims, level, batch):
    import torch

    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine
    from oracle import mlp as M
    from oracle import trainer as T

    prob = problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0)
    axes, P, h, bounds = grid_points(16)
    grid = DeviceGrid(*axes)
    rec = build_level(grid, T.level_spacing(h, level), bounds, problems.oracle_callbacks(prob), batch)
    theta = M.init_params(dims, seed=21)
    eager = Engine(dims, theta, OPT, max_rows=batch, max_steps=64)
    graphed = Engine(dims, theta, OPT, max_rows=batch, max_steps=64)
    assert eager.tile_path == (dims[1] != 64)  # tile kernels for 1x1/2x10, MFMA kernels for 3x64
    g = graphed.capture_epoch(grid, rec)
    assert int(graphed.count.item()) == 0  # capture records, nothing runs
    nb = len(rec.batch_begin)
    for epoch in range(3):
        eager.train_level(grid, rec, 1)
        g.replay()
        a, b = state(eager), state(graphed)
        for x, y in zip(a, b):
            assert torch.equal(x, y), epoch
        assert int(b[3].item()) == (epoch + 1) * nb
        # the program's own eager work between replays
        _ = float(graphed.loss_hist[: (epoch + 1) * nb].double().sum().item())
    hist = graphed.loss_hist[: 3 * nb].cpu().numpy()
    assert np.all(np.isfinite(hist)) and np.all(hist > 0)


@pytest.mark.gpu
def test_trainer_with_and_without_graphs(tmp_path):
    from jax_dips.solvers.poisson import trainer
    from oracle import mlp as M
    from test_gpu_trainer import run_trainer

    prob = problems.sphere()
    dims = (1, 1, 2, 10)
    theta0 = M.init_params(dims, seed=8)
    runs = []
    for use in (True, False):
        trainer.Trainer.use_graphs = use
        try:
            _, st, es, le, tr = run_trainer(prob, 16, 12, 1024, theta0, tmp_path / str(use))
        finally:
            trainer.Trainer.use_graphs = True
        runs.append((le, tr.model.flatten(tr.params), st.solution))
    (le_g, th_g, u_g), (le_e, th_e, u_e) = runs
    print("graph", le_g, "eager", le_e)
    assert np.array_equal(le_g, le_e)
    assert np.array_equal(th_g, th_e)
    assert np.array_equal(u_g, u_e)
