"""Learned per-point preconditioner — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates jax_dips/nn/preconditioner.py:10-35 and its use in
compute_Ax_and_b_preconditioned_fn (solvers/poisson/discretization.py:339,
:418-419): P(c) = 0.5 + s * sigmoid(MLP(c)) with MLP = [Dense(D_i), tanh]* then
Dense(1), evaluated on the 26 cell coefficients of the point
(geometric_integrations_per_point.py:873-900, :966-996: mu*A/h of the 12
(face, phase) pairs, V-, V+, and the 12 face areas); both lhs/diag and rhs/diag
are multiplied by P, so the residual becomes P * r0.

Parameters are one flat vector: for each Dense layer the kernel (fan_in x
fan_out, row-major, flax's [in, out] layout) then the bias.

Pinned by tests/test_oracle_precond.py: P is 0.5 exactly for s = 0 (a quarter
of the plain loss) and lies in (0.5, 0.5 + s); the VJP against central finite
differences of the preconditioned loss, for the network and the preconditioner
parameters.
"""
import numpy as np

N_COEFFS = 26


def shapes(widths):
    dims = [N_COEFFS] + list(widths) + [1]
    return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]


def size(widths):
    return sum(a * b + b for a, b in shapes(widths))


def unflatten(phi, widths):
    out, off = [], 0
    for a, b in shapes(widths):
        W = np.asarray(phi[off:off + a * b]).reshape(a, b)
        bb = np.asarray(phi[off + a * b:off + a * b + b])
        out.append((W, bb))
        off += a * b + b
    return out


def init_params(widths, seed=0, dt=np.float32):
    """Glorot-uniform kernels (preconditioner.py:22) and zero biases (flax Dense
    default); numpy's generator, not flax's PRNG stream."""
    rng = np.random.default_rng(seed)
    parts = []
    for a, b in shapes(widths):
        lim = np.sqrt(6.0 / (a + b))
        parts.append(rng.uniform(-lim, lim, size=a * b))
        parts.append(np.zeros(b))
    return np.concatenate(parts).astype(dt)


def coeffs26(slots, dt=np.float64):
    """The 26 cell coefficients of every point from oracle point_slots output."""
    geo = slots["geo"]
    n = slots["c_m"].shape[0]
    C = np.zeros((n, N_COEFFS), dtype=dt)
    C[:, 0:12:2] = slots["c_m"]
    C[:, 1:12:2] = slots["c_p"]
    C[:, 12] = geo["vol_m"]
    C[:, 13] = geo["vol_p"]
    C[:, 14:26:2] = geo["area_m"]
    C[:, 15:26:2] = geo["area_p"]
    return C


def forward(phi, widths, C, s, dt=np.float64):
    """P (n,) and the intermediate activations."""
    layers = unflatten(np.asarray(phi, dt), widths)
    acts = [np.asarray(C, dt)]
    a = acts[0]
    for W, b in layers[:-1]:
        a = np.tanh(a @ W.astype(dt) + b.astype(dt))
        acts.append(a)
    W, b = layers[-1]
    z = (a @ W.astype(dt) + b.astype(dt))[:, 0]
    sig = 1.0 / (1.0 + np.exp(-z))
    return dt(0.5) + dt(s) * sig, (acts, z, sig)


def vjp(phi, widths, C, s, seed_p, dt=np.float64):
    """Gradient of sum(seed_p * P) with respect to the flat parameters."""
    layers = unflatten(np.asarray(phi, dt), widths)
    _, (acts, z, sig) = forward(phi, widths, C, s, dt)
    d = (np.asarray(seed_p, dt) * dt(s) * sig * (1.0 - sig))[:, None]  # dL/dz
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, b = layers[li]
        a_in = acts[li]
        grads.append((a_in.T @ d, d.sum(axis=0)))
        if li > 0:
            d = (d @ W.astype(dt).T) * (1.0 - acts[li] ** 2)
    grads.reverse()
    return np.concatenate([np.concatenate([gW.ravel(), gb.ravel()]) for gW, gb in grads])
