"""The reference's per-step evaluation pattern on the CPU — the timed CPU baseline.

Test infrastructure only (see oracle/__init__.py); bench.py times it as the
``cpu_baseline`` (kind "port": a numpy restatement, jax is not installed).

What one training step costs in the reference (discretization.py:299-519,
trainer.py:783-789, 892-912): for every point of the batch, the geometry of the
7 stencil cells, the regression rows at crossed stencil points and every
callback are recomputed, and the network is evaluated at 7 x 27 cube
positions with BOTH phase networks (jnp.where evaluates both branches,
MLP.py:98), then differentiated.  ``step`` does exactly that work in
vectorised numpy (float32, like the reference) and returns the batch loss and
gradient; its loss equals the affine-record loss of oracle.trainer.
"""
import numpy as np

from . import discretization as D
from . import mlp as M


def step(theta, dims, P, h, bounds, cb, dt=np.float32, net=None):
    """Loss and gradient of one batch P (n,3), recomputing everything.  net:
    None for the two-phase MLP of ``dims``, else an object with
    ``forward(theta, X, phase_p, dt)`` and ``vjp(theta, X, phase_p, seed, dt)``
    (the hashgrid network, oracle/hashgrid.py)."""
    if net is not None:
        fwd, vjp = net.forward, net.vjp
    else:
        def fwd(th, X, ph, d):
            return M.forward(th, dims, X, ph, d)

        def vjp(th, X, ph, sd, d):
            return M.vjp(th, dims, X, ph, sd, d)
    slots = D.point_slots(P, h, bounds, cb, dt)
    W = slots["W"].reshape(P.shape[0], -1)  # (n, 189)
    pos = slots["pos"].reshape(-1, 3)
    phase = cb("phi", pos) >= 0
    # both networks at every slot, then the phase selection (jnp.where)
    up = fwd(theta, pos, np.ones(pos.shape[0], bool), dt)
    um = fwd(theta, pos, np.zeros(pos.shape[0], bool), dt)
    u = np.where(phase, up, um).reshape(W.shape)
    r = (W * u).sum(axis=1) + slots["c"]
    B = P.shape[0]
    loss = np.sum(dt(0.5) * r * r) / dt(B)
    seed = (r[:, None] * W / dt(B)).reshape(-1)
    # the VJP of jnp.where runs both branches (the unselected one with a zero cotangent)
    ones = np.ones(pos.shape[0], bool)
    grad = vjp(theta, pos, ones, seed * phase, dt) + vjp(theta, pos, ~ones, seed * ~phase, dt)
    return loss, grad
