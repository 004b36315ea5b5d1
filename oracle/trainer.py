"""Loss, gradient, batching, zoom schedule and epoch loop — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates:
  * Trainer.loss = mean 0.5 (lhs - rhs)^2 ........... trainer.py:892-912
  * Trainer.update (value_and_grad + optax) ....... trainer.py:783-789
  * single_GPU_train (scan epochs x batches; loss_epochs[e] = mean of the
    pre-update batch losses; no shuffling) ........ trainer.py:501-591
  * alternate_res_sequentially: h = h_tr * 0.5**(epoch // (E // 4))
                                                  data_management.py:319-326
  * DatasetDict contiguous batches (+ padding when B does not divide N)
                                                  data_management.py:79-160

The per-point residual comes from oracle.discretization in affine form; the
network is evaluated at the canonical float32 lattice positions fl32(p + o h).
Padding points (B not dividing N) use numpy's generator instead of jax's
PRNGKey(0) normals (same rescaling rule), so their coordinates differ from the
reference's; none of the benchmark workloads pads.
"""
import numpy as np

from . import discretization as D
from . import mlp as M
from . import optim as O
from . import precond as PC


def zoom_level(epoch, num_epochs):
    return epoch // (num_epochs // 4)


def level_spacing(h_tr, level):
    return (np.asarray(h_tr, dtype=np.float32) * np.float32(0.5) ** np.float32(level)).astype(np.float32)


def batches(n, batch_size):
    """Index lists of contiguous batches; last one padded with -1 entries."""
    B = min(batch_size, n)
    nb = -(-n // B)
    out = []
    for b in range(nb):
        idx = np.arange(b * B, min((b + 1) * B, n))
        if idx.size < B:
            idx = np.concatenate([idx, -np.ones(B - idx.size, dtype=np.int64)])
        out.append(idx)
    return out


def padding_points(R, count, seed=0):
    """generate_random_points_like (data_management.py:70-76), numpy generator."""
    rng = np.random.default_rng(seed)
    dR = rng.standard_normal((count, 3))
    lo, hi = R.min(), R.max()
    return (lo + (hi - lo) * (dR - dR.min()) / (dR.max() - dR.min())).astype(np.float32)


class LevelRecords:
    """Dense (offset, phase)-keyed affine records of all points at one stencil
    spacing: r = c + sum_k W[:, k] * NN_{k % 2}(pos[:, k // 2])."""

    def __init__(self, P, h, bounds, cb, dt=np.float64):
        slots = D.point_slots(P, h, bounds, cb, dt)
        self.W, self.pos = D.merge_slots(slots, P, h, lambda R: cb("phi", R))
        self.c = slots["c"]
        self.boundary = slots["boundary"]
        self.dead = slots["dead"]
        self.slots = slots
        self.h = np.asarray(h, np.float32)

    @property
    def coeffs(self):
        """The 26 cell coefficients of every point (preconditioner input)."""
        if getattr(self, "_coeffs", None) is None:
            self._coeffs = PC.coeffs26(self.slots)
        return self._coeffs

    def entries(self, rows):
        W = self.W[rows]
        nz = np.nonzero(W)
        X = self.pos[rows][nz[0], nz[1] // 2]
        ph = (nz[1] % 2) == 1
        return W, nz, X, ph


def residual(theta, dims, rec, rows, dt=np.float64):
    """r for the given rows (n,), evaluating the network only where W != 0."""
    W, nz, X, ph = rec.entries(rows)
    u = M.forward(theta, dims, X, ph, dt)
    r = rec.c[rows].astype(dt).copy()
    np.add.at(r, nz[0], W[nz].astype(dt) * u)
    return r


def loss_and_grad(theta, dims, rec, rows, dt=np.float64, precond=None, act="tanh"):
    """0.5 * mean(r^2) over the batch and its gradient (hand VJP).

    precond = dict(widths=..., scale=s): the last precond.size(widths) entries
    of theta are the preconditioner's, r = P * r0 with P from the 26 cell
    coefficients of each point (discretization.py:339, :418-419); the gradient
    covers both parameter sets."""
    W, nz, X, ph = rec.entries(rows)
    nn = M.layout(dims)["size"]
    th = np.asarray(theta)[:nn]
    u = M.forward(th, dims, X, ph, dt, act)
    r0 = rec.c[rows].astype(dt).copy()
    np.add.at(r0, nz[0], W[nz].astype(dt) * u)
    B = len(rows)
    if precond is None:
        loss = np.sum(dt(0.5) * r0 * r0) / dt(B)
        seed = r0[nz[0]] * W[nz].astype(dt) / dt(B)
        return loss, M.vjp(th, dims, X, ph, seed, dt, act), r0
    phi = np.asarray(theta)[nn:]
    C = rec.coeffs[rows]
    P, _ = PC.forward(phi, precond["widths"], C, precond["scale"], dt)
    r = P * r0
    loss = np.sum(dt(0.5) * r * r) / dt(B)
    seed = (P * P * r0)[nz[0]] * W[nz].astype(dt) / dt(B)
    g_nn = M.vjp(th, dims, X, ph, seed, dt, act)
    g_pc = PC.vjp(phi, precond["widths"], C, precond["scale"], P * r0 * r0 / dt(B), dt)
    return loss, np.concatenate([g_nn, g_pc]), r


def train(P, h_tr, bounds, cb, dims, theta0, num_epochs, batch_size, opt_cfg,
          dt=np.float64, param_dt=np.float32, record_dt=np.float64, records=None, precond=None, act="tanh",
          loss_fn=None):
    """Epoch loop of single_GPU_train with the zoom schedule.  Returns
    (theta, loss_epochs, per-step losses).  A short last batch is padded with
    padding_points (DatasetDict.gpu_padded_batches, data_management.py:103-173),
    which become training points after the grid points.  loss_fn(theta, rec,
    rows) -> (loss, grad, r) replaces the two-phase MLP's loss_and_grad (the
    hashgrid network, oracle/hashgrid.py)."""
    n = P.shape[0]
    bl = batches(n, batch_size)
    npad = int(sum((b < 0).sum() for b in bl))
    if npad:
        base = P
        P = np.concatenate([P, padding_points(base, npad)]).astype(np.float32)
        bl = [np.arange(k * len(bl[0]), (k + 1) * len(bl[0])) for k in range(len(bl))]
    theta = np.asarray(theta0, dtype=param_dt).copy()
    state = O.OptState(theta.size, param_dt)
    recs = {} if records is None else records
    loss_epochs = np.zeros(num_epochs)
    steps = []
    for epoch in range(num_epochs):
        lvl = zoom_level(epoch, num_epochs)
        if lvl not in recs:
            recs[lvl] = LevelRecords(P, level_spacing(h_tr, lvl), bounds, cb, record_dt)
        rec = recs[lvl]
        tot = 0.0
        for rows in bl:
            if loss_fn is not None:
                loss, grad, _ = loss_fn(theta, rec, rows)
            else:
                loss, grad, _ = loss_and_grad(theta, dims, rec, rows, dt, precond, act)
            theta = O.step(theta, grad.astype(param_dt), state, opt_cfg, param_dt)
            tot += loss
            steps.append(loss)
        loss_epochs[epoch] = tot / len(bl)
    return theta, loss_epochs, np.array(steps), recs
