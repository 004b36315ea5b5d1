"""GPU parity of the packed-tile training step (dips_tile_*) against the CPU oracle.

Records come from the oracle (float64) so the test isolates the tile kernel:
one lane per entry, segmented row sums, MFMA-accumulated weight gradients for
networks of width 8..12 and per-lane gradients for the small ones, plus the
fused last-workgroup optimizer.  Tolerances as in test_gpu_loss.py: float32
network and sums against the float64 oracle give 1e-5 relative on the loss and
1e-4 relative (2-norm) on the gradient.
"""
import numpy as np
import pytest

from helpers import oracle_level
from jax_dips import problems
from oracle import mlp as M
from oracle import optim as O
from oracle import trainer as T

DIMS = (1, 1, 2, 10)


def _device(axes, rec, batch_size):
    from jax_dips.records import DeviceGrid, DeviceRecords, pack_dense

    packed = pack_dense(rec.W, rec.c, batch_size)
    return DeviceGrid(*axes), DeviceRecords(packed, rec.h), packed


def _check(loss, grad, ol, og, what):
    print(f"{what}: loss gpu {loss:.9e} oracle {ol:.9e} grad rel "
          f"{np.linalg.norm(grad - og) / np.linalg.norm(og):.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol), what
    assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og), what


@pytest.mark.gpu
@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_tile_loss_grad_sphere16(level):
    from jax_dips.records import Engine

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    n = P.shape[0]
    theta = M.init_params(DIMS, seed=3)
    grid, drec, _ = _device(axes, rec, n)
    eng = Engine(DIMS, theta, {}, max_rows=n, max_steps=4)
    assert eng.tile_path
    loss, grad = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), DIMS, rec, np.arange(n))
    _check(loss, grad, ol, og, f"level {level}")


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(2, 10, 2, 10), (1, 3, 2, 10), (1, 1, 1, 1), (2, 16, 2, 16)])
def test_tile_loss_grad_shapes(dims):
    """MFMA accumulation in both networks (width 16: weight blocks on the matrix
    cores, biases and output layer per lane), and per-lane accumulation only."""
    from jax_dips.records import Engine

    prob = problems.sphere(mu_m=2.0, mu_p=7.0, k_m=0.5, k_p=1.5)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    n = P.shape[0]
    theta = M.init_params(dims, seed=11)
    grid, drec, _ = _device(axes, rec, n)
    eng = Engine(dims, theta, {}, max_rows=n, max_steps=4)
    loss, grad = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, rec, np.arange(n))
    _check(loss, grad, ol, og, str(dims))


def _synthetic_long_rows(rec, seed):
    """The records with some rows widened to 70..250 entries (random weights on
    extra (offset, phase) slots), some emptied (constant only), to exercise the
    one-wavefront-per-row path and rows without entries."""
    rng = np.random.default_rng(seed)
    W = rec.W.copy()
    n = W.shape[0]
    rows = rng.choice(n, 40, replace=False)
    for i, r in enumerate(rows[:30]):
        k = [70, 100, 130, 250][i % 4]
        slots = rng.choice(250, k, replace=False)
        W[r, slots] = rng.normal(0.0, 0.05, k)
    W[rows[30:]] = 0.0
    out = T.LevelRecords.__new__(T.LevelRecords)
    out.W, out.pos, out.c, out.h = W, rec.pos, rec.c, rec.h
    return out, rows


@pytest.mark.gpu
def test_tile_long_and_empty_rows():
    from jax_dips.records import Engine

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec0 = oracle_level(prob, 16, 2)
    rec, rows = _synthetic_long_rows(rec0, 5)
    n = P.shape[0]
    B = 1024
    theta = M.init_params(DIMS, seed=17)
    grid, drec, _ = _device(axes, rec, B)
    tl = drec.tiles(grid)
    ptr = drec.row_ptr.cpu().numpy().astype(np.int64)
    trow = tl["tile_row"].cpu().numpy().astype(np.int64)
    long_tiles = sum(1 for b in range(len(drec.batch_begin))
                     for t in range(tl["batch_tile_begin"][b], tl["batch_tile_begin"][b + 1] - 1)
                     if ptr[trow[t + 1]] - ptr[trow[t]] > 64)
    assert long_tiles >= 20
    eng = Engine(DIMS, theta, {}, max_rows=B, max_steps=4)
    for b, (b0, nr) in enumerate(zip(drec.batch_begin, drec.batch_rows)):
        loss, grad = eng.tile_loss_grad(grid, drec, b, 1.0 / B)
        ol, og, _ = T.loss_and_grad(theta.astype(np.float64), DIMS, rec, np.arange(b0, b0 + nr))
        _check(loss, grad, ol * nr / B, og * nr / B, f"batch {b}")


@pytest.mark.gpu
def test_tile_batches_uneven():
    """Batches of 1000 rows over 20^3 points (last batch shorter)."""
    from jax_dips.records import Engine

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 20, 0)
    B = 1000
    theta = M.init_params(DIMS, seed=5)
    grid, drec, _ = _device(axes, rec, B)
    eng = Engine(DIMS, theta, {}, max_rows=B, max_steps=4)
    for b, (b0, nr) in enumerate(zip(drec.batch_begin, drec.batch_rows)):
        loss, grad = eng.tile_loss_grad(grid, drec, b, 1.0 / B)
        ol, og, _ = T.loss_and_grad(theta.astype(np.float64), DIMS, rec, np.arange(b0, b0 + nr))
        _check(loss, grad, ol * nr / B, og * nr / B, f"batch {b}")


@pytest.mark.gpu
def test_tile_train_level_matches_oracle_steps():
    """20 fused tile steps (5 epochs x 4 batches) against the oracle loop with
    the float32 optax restatement."""
    from jax_dips.records import Engine
    import torch

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 0)
    n = P.shape[0]
    B = 1024
    theta0 = M.init_params(DIMS, seed=7)
    grid, drec, _ = _device(axes, rec, B)
    cfg = {"optimizer_name": "custom", "learning_rate": 1e-2,
           "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    eng = Engine(DIMS, theta0, cfg, max_rows=B, max_steps=64)
    assert eng.tile_path and eng.use_tiles
    eng.train_level(grid, drec, 5)
    torch.cuda.synchronize()
    gl = eng.loss_hist[:20].cpu().numpy().astype(np.float64)
    gp = eng.params.cpu().numpy()
    th = theta0.copy()
    st = O.OptState(th.size, np.float32)
    ol = []
    for e in range(5):
        for b0 in range(0, n, B):
            l, g, _ = T.loss_and_grad(th.astype(np.float64), DIMS, rec, np.arange(b0, b0 + B))
            ol.append(l)
            th = O.step(th, g.astype(np.float32), st, cfg, np.float32)
    ol = np.array(ol)
    print("gpu", gl[:4], "oracle", ol[:4])
    assert np.all(np.abs(gl - ol) <= 1e-4 * np.abs(ol))
    assert np.linalg.norm(gp - th) <= 1e-3 * np.linalg.norm(th)
    assert int(eng.count.item()) == 20 and int(eng.sync.item()) == 0


@pytest.mark.gpu
def test_tile_fused_equals_two_kernel_and_row_kernel():
    """The single-launch tile step (last workgroup applies the optimizer) equals
    the two-launch step bit for bit in the losses, and tracks the per-row
    kernel's trajectory."""
    from jax_dips.records import Engine
    import torch

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    B = 1024
    theta0 = M.init_params(DIMS, seed=13)
    grid, drec, _ = _device(axes, rec, B)
    cfg = {"optimizer_name": "custom", "learning_rate": 1e-2,
           "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    out = []
    for single, tiles in ((True, True), (False, True), (True, False)):
        eng = Engine(DIMS, theta0, cfg, max_rows=B, max_steps=64)
        eng.tile_single_launch, eng.use_tiles = single, tiles
        eng.train_level(grid, drec, 6)
        torch.cuda.synchronize()
        out.append((eng.loss_hist[:24].cpu().numpy(), eng.params.cpu().numpy(), int(eng.count.item())))
    (l1, p1, c1), (l2, p2, c2), (l3, p3, c3) = out
    assert c1 == c2 == c3 == 24
    assert np.array_equal(l1, l2)
    assert np.max(np.abs(p1 - p2)) <= 1e-7 * max(1.0, np.max(np.abs(p2)))
    assert np.all(np.abs(l1 - l3) <= 1e-5 * np.abs(l3))


@pytest.mark.gpu
@pytest.mark.parametrize("G", [2, 8])
def test_fake_data_parallel_shards(G):
    """One-GPU rehearsal of data parallelism: a batch's tiles split into G
    contiguous shards (what G ranks would each take), the tile kernel run per
    shard, the shard gradients summed (the all-reduce) equal the full batch's
    gradient; the shards are disjoint, so the sums differ only by float32
    summation order (1e-6 relative)."""
    import ctypes

    from jax_dips import _native as N
    from jax_dips.records import Engine
    import torch

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 20, 1)
    n = P.shape[0]
    theta = M.init_params(DIMS, seed=31)
    grid, drec, _ = _device(axes, rec, n)
    eng = Engine(DIMS, theta, {}, max_rows=n, max_steps=4)
    full_l, full_g = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    tl = drec.tiles(grid)
    tb = int(tl["batch_tile_begin"][0])
    nt = int(tl["batch_tile_begin"][1]) - tb - 1
    cuts = np.linspace(0, nt, G + 1).astype(int)
    tot_l, tot_g = 0.0, np.zeros_like(full_g)
    for k in range(G):
        nblk = ctypes.c_int32(0)
        N.check(N.lib().dips_tile_loss_grad(ctypes.byref(grid.c), ctypes.byref(drec.c), ctypes.byref(tl["c"]),
                                            tb + int(cuts[k]), int(cuts[k + 1] - cuts[k]), ctypes.byref(eng.net),
                                            N.ptr(eng.params), N.ptr(eng.partial), eng.partial_cap, 1,
                                            ctypes.byref(nblk), N.current_stream()), "dips_tile_loss_grad")
        torch.cuda.synchronize()
        part = eng.partial[: nblk.value * (eng.P + 1)].view(nblk.value, eng.P + 1).double().cpu().numpy()
        s = part.sum(axis=0) / n
        tot_l += s[eng.P]
        tot_g += s[: eng.P]
    assert abs(tot_l - full_l) <= 1e-6 * abs(full_l)
    assert np.linalg.norm(tot_g - full_g) <= 1e-6 * np.linalg.norm(full_g)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), DIMS, rec, np.arange(n))
    _check(tot_l, tot_g, ol, og, f"{G} shards")
