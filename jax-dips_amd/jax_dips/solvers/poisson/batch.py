"""Batched training of many independent interface problems on one GPU.

The reference trains one problem per trainer run and sweeps several problems
as a pool of processes, one per GPU (examples/biomolecules/main.py:400-432,
each process running solvers/poisson/trainer.py).  ``solve_batch`` trains all
problems of the sweep together: per zoom level the FP64 geometry kernels build
each problem's records (level_records.build_level), then dips_batch_train runs
every epoch of the level for all problems, one step launch over the current
batch of every problem followed by one reduction + optimizer workgroup per
problem.  Each problem's losses and parameters are exactly those of a
single-problem ``Trainer`` run with the same initial parameters (the same
kernels, block counts and reduction order).

Problems share the network shape, optimizer settings and epoch count; grid
sizes, batch sizes (default: min(N, batch_size)), callbacks and initial
parameters are per problem.  A short last batch is padded with random points
in the data's range, as Trainer.solve does.
"""
import logging

import numpy as np

from ...data.data_management import TrainData, generate_random_points_like
from ...level_records import build_level, sample
from ...nn.configure import get_model
from ...nn.preconditioner import Preconditioner
from ...records import BatchEngine, DeviceGrid
from .trainer import DEFAULT_MODEL, DEFAULT_OPTIMIZER

logger = logging.getLogger(__name__)

CALLBACKS = {"phi": "lvl_set_fn", "mu_m": "mu_m_fn", "mu_p": "mu_p_fn", "k_m": "k_m_fn", "k_p": "k_p_fn",
             "f_m": "f_m_fn", "f_p": "f_p_fn", "alpha": "alpha_fn", "beta": "beta_fn", "g": "dirichlet_bc_fn"}


def _flatten(model, precond, tree):
    flat = model.flatten(tree)
    if precond is not None:
        flat = np.concatenate([flat, precond.flatten(tree["preconditioner"])])
    return flat


def _unflatten(model, precond, flat):
    nn = sum(a * b + b for tag in ("p", "m") for a, b in model.shapes(tag))
    pre = precond.unflatten(flat[nn:]) if precond is not None else {}
    return dict(model.unflatten(flat[:nn]), preconditioner=pre)


def solve_batch(problems, num_epochs=1000, batch_size=65536, model_dict=DEFAULT_MODEL,
                optimizer_dict=DEFAULT_OPTIMIZER, init_params=None, evaluate=True):
    """Train every problem of ``problems`` for ``num_epochs`` epochs.

    problems: list of dicts holding the callbacks of ``trainer.setup``
    (``lvl_set_fn``, ``dirichlet_bc_fn``, ``mu_m_fn``, ``mu_p_fn``, ``k_m_fn``,
    ``k_p_fn``, ``f_m_fn``, ``f_p_fn``, ``alpha_fn``, ``beta_fn``), the grid
    states ``lvl_gstate`` and ``tr_gstate``, optionally ``eval_gstate`` and
    ``batch_size``.  init_params: None, or one flat vector / params tree per
    problem.  Returns one dict per problem: ``params`` (tree),
    ``loss_epochs`` (float32[E]) and, with ``evaluate`` and an eval grid,
    ``solution`` on it.
    """
    import torch

    if not problems:
        return []
    E = int(num_epochs)
    if E < 4:
        raise ValueError("num_epochs must be at least 4 (four zoom levels of E // 4 epochs)")
    forward, _ = get_model(model_dict, model_type=model_dict.get("model_type", "mlp"))
    model = forward.model
    if getattr(model, "encoder", "identity") in ("hashgrid", "frequency", "spherical_harmonics"):
        raise NotImplementedError("batched training of the hashgrid / frequency encoders has no device "
                                  "implementation; train each problem with trainer.setup")
    pc = model_dict.get("preconditioner") or {}
    precond = Preconditioner(pc["layer_widths"], 1, pc.get("scaling_coeff", 1.0)) if pc.get("enable", False) else None

    grids, tds, bounds, cbs, nbs, params0 = [], [], [], [], [], []
    for i, prob in enumerate(problems):
        tr = prob["tr_gstate"]
        n = len(tr.R)
        B = min(int(prob.get("batch_size", batch_size)), n)
        nb = -(-n // B)
        extra = generate_random_points_like(np.asarray(tr.R), nb * B - n) if nb * B > n else None
        grids.append(DeviceGrid(tr.x, tr.y, tr.z, extra=extra))
        tds.append((TrainData(tr, prob["lvl_set_fn"]), B))
        lv = prob["lvl_gstate"]
        bounds.append((lv.xmin(), lv.xmax(), lv.ymin(), lv.ymax(), lv.zmin(), lv.zmax()))
        cbs.append({k: prob[v] for k, v in CALLBACKS.items()})
        nbs.append(nb)
        if init_params is None:
            tree = dict(forward.init(42 + i))
            if precond is not None:
                tree["preconditioner"] = precond.init(42 + i)
            params0.append(_flatten(model, precond, tree))
        else:
            p = init_params[i]
            params0.append(_flatten(model, precond, p) if isinstance(p, dict) else np.asarray(p, np.float32))
    eng = BatchEngine(model.dims, params0, optimizer_dict, [E * nb for nb in nbs],
                      precond=precond.engine_config() if precond is not None else None,
                      act=getattr(model, "act", 0))
    epoch = 0
    while epoch < E:
        lvl = epoch // (E // 4)
        last = epoch
        while last + 1 < E and (last + 1) // (E // 4) == lvl:
            last += 1
        recs = []
        for i, prob in enumerate(problems):
            td, B = tds[i]
            tr = prob["tr_gstate"]
            h = np.array(td.alternate_res_sequentially(E, epoch, tr.dx, tr.dy, tr.dz), np.float32)
            recs.append(build_level(grids[i], h, bounds[i], cbs[i], B, coeffs=precond is not None))
        torch.cuda.synchronize()
        eng.train_level(grids, recs, last + 1 - epoch)
        torch.cuda.synchronize()
        del recs
        logger.info(f"batch: epochs {epoch}..{last} (level {lvl}) done for {len(problems)} problems")
        epoch = last + 1

    out = []
    for i, (prob, e) in enumerate(zip(problems, eng.engines)):
        hist = e.loss_hist.cpu().numpy()
        loss_epochs = hist[:E * nbs[i]].reshape(E, nbs[i]).mean(axis=1).astype(np.float32)
        flat = e.params.cpu().numpy()
        res = {"params": _unflatten(model, precond, flat), "loss_epochs": loss_epochs}
        ev = prob.get("eval_gstate")
        if evaluate and ev is not None:
            from ...evaluation import evaluate_points

            R = np.asarray(ev.R, np.float32)
            phase = sample(cbs[i]["phi"], torch.as_tensor(R, device="cuda")).cpu().numpy() >= 0
            res["solution"], _ = evaluate_points(model, flat[:model.flatten(model.unflatten(flat)).size], R, phase)
        out.append(res)
    return out
