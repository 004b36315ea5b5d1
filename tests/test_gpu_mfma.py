"""GPU parity of the MFMA training-step kernels (wide networks) with the CPU
oracle: loss and mean gradient of a batch, and optimizer trajectories.

Tolerance: float32 accumulation over hidden widths up to 64 (the forward
pass's hidden layers from width 32 on as exact three-way bf16 splits on the
bf16 matrix cores, every other contraction on the float32 ones) against the
float64 oracle: loss within 1e-5 relative, gradient within 1e-4 relative in the
2-norm (as for the lane-per-row kernel).
"""
import numpy as np
import pytest

from helpers import oracle_level
from jax_dips import problems
from oracle import mlp as M
from oracle import optim as O
from oracle import trainer as T


def device(axes, rec, B):
    from jax_dips.records import DeviceGrid, DeviceRecords, pack_dense

    return DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h)


@pytest.mark.gpu
@pytest.mark.parametrize("dims,level,grouped", [((2, 16, 2, 16), 0, True), ((2, 16, 2, 16), 2, True),
                                                ((3, 64, 3, 64), 1, True), ((3, 64, 3, 64), 0, False),
                                                ((3, 64, 3, 64), 0, True), ((1, 1, 2, 10), 0, True),
                                                ((3, 32, 3, 32), 3, True), ((3, 32, 3, 32), 1, False)])
def test_mfma_loss_grad(dims, level, grouped):
    """grouped: record entries at one position evaluated once (their u written
    to every member, their seeds summed); else every entry on its own."""
    from jax_dips.records import Engine

    prob = problems.star(r0=0.55, amp=0.15, centre=(0.0123, -0.0171, 0.0093))
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    n = P.shape[0]
    B = 2048
    theta = M.init_params(dims, seed=31)
    grid, drec = device(axes, rec, B)
    eng = Engine(dims, theta, {}, max_rows=B, max_steps=4)
    if grouped and level <= 1:
        mb = drec.mfma_batches(grid)
        assert mb["groups"] < drec.nnz  # at zoom levels 0 and 1 stencil neighbours share positions
    for b in range(n // B):
        loss, grad = eng.mfma_loss_grad(grid, drec, b, 1.0 / B, grouped=grouped)
        ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, rec, np.arange(b * B, (b + 1) * B))
        print(dims, level, b, "loss", loss, ol, "grad rel", np.linalg.norm(grad - og) / np.linalg.norm(og))
        assert abs(loss - ol) <= 1e-5 * abs(ol)
        assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
def test_mfma_training_steps_match_oracle():
    from jax_dips.records import Engine
    import torch

    dims = (2, 16, 2, 16)
    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    B = 1024
    theta0 = M.init_params(dims, seed=5)
    grid, drec = device(axes, rec, B)
    cfg = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    eng = Engine(dims, theta0, cfg, max_rows=B, max_steps=32)
    eng.train_level_mfma(grid, drec, 3)
    torch.cuda.synchronize()
    gl = eng.loss_hist[:12].cpu().numpy().astype(np.float64)
    th = theta0.copy()
    st = O.OptState(th.size, np.float32)
    ol = []
    for e in range(3):
        for b0 in range(0, P.shape[0], B):
            l, g, _ = T.loss_and_grad(th.astype(np.float64), dims, rec, np.arange(b0, b0 + B))
            ol.append(l)
            th = O.step(th, g.astype(np.float32), st, cfg, np.float32)
    ol = np.array(ol)
    assert int(eng.count.item()) == 12
    assert np.all(np.abs(gl - ol) <= 1e-4 * np.abs(ol))
    assert np.linalg.norm(eng.params.cpu().numpy() - th) <= 1e-3 * np.linalg.norm(th)


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(3, 64, 3, 64), (2, 16, 2, 16), (1, 1, 3, 64), (3, 32, 3, 32)])
def test_stored_and_recomputed_activations_agree(dims):
    """The backward kernel reads the forward tiles' stored activations when the
    workspace holds them and recomputes them otherwise; both give the same bits
    (loss and gradient), and the workspace-size query covers the stored form."""
    import ctypes

    import torch

    from jax_dips import _native as N
    from jax_dips.records import Engine

    prob = problems.star(r0=0.55, amp=0.15, centre=(0.0123, -0.0171, 0.0093))
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 0)
    B = 2048
    theta = M.init_params(dims, seed=7)
    grid, drec = device(axes, rec, B)
    eng = Engine(dims, theta, {}, max_rows=B, max_steps=4)
    mb = eng.mfma_workspace(drec, grid)
    L = N.lib()
    full = int(eng._mwork.numel())
    # the same workspace without its activation region: 64 alignment floats and max_nnz / 16 + 2 tiles of
    # max((L - 1) x ceil(H / 16)) f4s per lane (hidden layers 1..L-1)
    act_f4 = max((dims[0] - 1) * ((dims[1] + 15) // 16), (dims[2] - 1) * ((dims[3] + 15) // 16))
    base = full - 64 - (mb["max_nnz"] // 16 + 2) * act_f4 * 256
    assert base > 0
    for b in range(min(2, P.shape[0] // B)):
        out = []
        for wf in (full, base):
            g = torch.empty(eng.P, dtype=torch.float32, device=eng.device)
            loss = torch.empty(1, dtype=torch.float64, device=eng.device)
            N.check(L.dips_mfma_loss_grad(ctypes.byref(grid.c), ctypes.byref(drec.c), ctypes.byref(mb["batches"][b]),
                                          ctypes.byref(eng.net), ctypes.c_float(1.0 / B), N.ptr(eng.params), N.ptr(g),
                                          N.ptr(loss), N.ptr(eng._mwork), wf, N.ptr(eng._mdwork),
                                          int(eng._mdwork.numel()), N.current_stream()), "dips_mfma_loss_grad")
            torch.cuda.synchronize()
            out.append((loss.clone(), g.clone()))
        print(dims, b, "loss", float(out[0][0]), float(out[1][0]))
        assert torch.equal(out[0][0], out[1][0])
        assert torch.equal(out[0][1], out[1][1])
