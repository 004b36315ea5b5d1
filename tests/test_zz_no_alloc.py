"""The allocation-free MFMA step, dips_mfma_train_step_noalloc (workspace in
the library's fixed per-device scratch array), against dips_mfma_train_step
with a caller-given workspace: bit-identical parameters, Adam moments, step
counter and loss history after each of two calls in a row, the second on a
smaller batch (fewer rows and entries, so nothing the first call left in the
shared array may reach the second's results).  Oversized batches are
rejected on the host before any device work."""
import ctypes

import numpy as np
import pytest

from helpers import grid_points
from jax_dips import problems

OPT = {"optimizer_name": "custom", "learning_rate": 1e-3, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


@pytest.mark.gpu
def test_noalloc_step_equals_workspace_step():
    import torch

    from jax_dips import _native as N
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine
    from oracle import mlp as M

    prob = problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0)
    axes, P, h, bounds = grid_points(24)
    grid = DeviceGrid(*axes)
    cbs = problems.oracle_callbacks(prob)
    big = build_level(grid, h, bounds, cbs, 8192)
    small = build_level(grid, (h * np.float32(0.5)).astype(np.float32), bounds, cbs, 2048)
    dims = (3, 64, 3, 64)
    theta = M.init_params(dims, seed=3)
    a = Engine(dims, theta, OPT, max_rows=8192, max_steps=4)
    b = Engine(dims, theta, OPT, max_rows=8192, max_steps=4)
    L = N.lib()
    stream = N.current_stream()
    sizes = []
    for rec, bi in ((big, 0), (small, 1)):
        mb = a.mfma_workspace(rec, grid)
        bt = mb["batches"][bi]
        sizes.append((bt.nrows, bt.ent_end - bt.ent_begin))
        inv_b = ctypes.c_float(1.0 / rec.batch_size)
        N.check(L.dips_mfma_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(bt),
                                       ctypes.byref(a.net), ctypes.byref(a.opt), inv_b, N.ptr(a.params), N.ptr(a.m),
                                       N.ptr(a.v), N.ptr(a.count), N.ptr(a.loss_hist), int(a.loss_hist.numel()),
                                       N.ptr(a._mwork), int(a._mwork.numel()), N.ptr(a._mdwork),
                                       int(a._mdwork.numel()), stream), "dips_mfma_train_step")
        N.check(L.dips_mfma_train_step_noalloc(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(bt),
                                               ctypes.byref(b.net), ctypes.byref(b.opt), inv_b, N.ptr(b.params),
                                               N.ptr(b.m), N.ptr(b.v), N.ptr(b.count), N.ptr(b.loss_hist),
                                               int(b.loss_hist.numel()), stream), "dips_mfma_train_step_noalloc")
        torch.cuda.synchronize()
        for x, y in ((a.params, b.params), (a.m, b.m), (a.v, b.v), (a.count, b.count), (a.loss_hist, b.loss_hist)):
            assert torch.equal(x, y)
    print("batches (rows, entries):", sizes, "losses:", b.loss_hist[:2].cpu().numpy())
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]
    assert int(b.count.item()) == 2
    assert np.all(np.isfinite(b.loss_hist[:2].cpu().numpy())) and float(b.loss_hist[1]) > 0


def test_noalloc_rejects_oversized_batch():
    """A batch whose workspace exceeds the fixed scratch array returns
    DIPS_EINVAL with a message naming the limit; the check runs on the host
    before any device call (the pointers here are never dereferenced)."""
    from jax_dips import _native as N

    L = N.lib()
    fake = 256
    grid = N.Grid(fake, fake, fake, 256, 256, 256, None, 0)
    rows, nnz = 1 << 20, 8 << 20
    rec = N.Records(fake, fake, fake, fake, fake, rows, nnz, 0.01, 0.01, 0.01, None)
    bt = N.MfmaBatch(0, rows, 0, nnz, fake, fake, fake, fake, 0, nnz // 2, nnz - nnz // 2, nnz)
    net = N.Net(3, 64, 3, 64, 0, 0, 0.0, 0)
    opt = N.Opt(0, 0, 1e-3, 0.9, 1000.0, 1.0)
    st = L.dips_mfma_train_step_noalloc(ctypes.byref(grid), ctypes.byref(rec), ctypes.byref(bt), ctypes.byref(net),
                                        ctypes.byref(opt), ctypes.c_float(1.0 / rows), fake, fake, fake, fake, fake,
                                        16, None)
    assert st == 1
    assert "beyond the fixed" in L.dips_last_error().decode()
