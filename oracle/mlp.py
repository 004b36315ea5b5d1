"""Two-phase tanh MLP (DoubleMLP) forward and hand-written VJP — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates jax_dips/nn/mlp/MLP.py:93-139: u(x) = NN+(x) if phi(x) >= 0 else
NN-(x); NN = [Linear(H), tanh] * L then Linear(1).  Parameters are handled as
one flat vector laid out as (for net in (+, -)): for each layer W (fan_in x
fan_out, row-major) then b.  ``dims`` = (L_m, H_m, L_p, H_p).

act = "celu" restates the hidden activation of the coordinate MLPs of the
multiresolution hash network (nn/hash_encoding/blocks/nerfs_flax.py:142-170,
flax nn.celu with alpha = 1: x for x > 0, expm1(x) otherwise).

Pinned by tests/test_oracle_mlp.py: forward against an independent torch
(float64, CPU) implementation; the VJP against torch autograd and against
float64 central finite differences (h = 1e-3 on 32 random parameters), for
both activations; CELU against its closed form.
"""
import numpy as np


def activation(h, act):
    if act == "tanh":
        return np.tanh(h)
    if act == "celu":
        return np.where(h > 0, h, np.expm1(np.minimum(h, 0.0)))
    raise ValueError(act)


def activation_grad(a, act):
    """Derivative from the activation value a (tanh: 1 - a^2; celu: 1 above 0, a + 1 below)."""
    if act == "tanh":
        return 1 - a * a
    return np.where(a > 0, 1.0, a + 1.0)


def net_shapes(L, H):
    shapes = [(3, H)] + [(H, H)] * (L - 1) + [(H, 1)]
    return shapes


def net_size(L, H):
    return sum(a * b + b for a, b in net_shapes(L, H))


def layout(dims):
    """Offsets of every (W, b) in the flat vector: list for the + net then the - net."""
    Lm, Hm, Lp, Hp = dims
    out = {}
    off = 0
    for tag, (L, H) in (("p", (Lp, Hp)), ("m", (Lm, Hm))):
        lay = []
        for a, b in net_shapes(L, H):
            lay.append((off, a, b))
            off += a * b + b
        out[tag] = lay
    out["size"] = off
    return out


def unflatten(theta, dims):
    lay = layout(dims)
    nets = {}
    for tag in ("p", "m"):
        nets[tag] = []
        for off, a, b in lay[tag]:
            W = theta[off:off + a * b].reshape(a, b)
            bb = theta[off + a * b:off + a * b + b]
            nets[tag].append((W, bb))
    return nets


def init_params(dims, seed=42, dt=np.float32):
    """Haiku-style initialisation (MLP.py:65, 70, 112, 115, 129, 138): hidden
    weights TruncatedNormal(stddev 0.1), output weights TruncatedNormal(1/sqrt(fan_in)),
    biases zero.  Uses numpy's generator, so the draws differ from haiku's
    PRNGKey(42) stream; callers that need a specific start pass explicit params."""
    rng = np.random.default_rng(seed)
    lay = layout(dims)
    theta = np.zeros(lay["size"], dtype=np.float64)
    for tag in ("p", "m"):
        n = len(lay[tag])
        for i, (off, a, b) in enumerate(lay[tag]):
            std = 0.1 if i < n - 1 else 1.0 / np.sqrt(a)
            real = std / 0.87962566103423978
            w = rng.standard_normal(a * b)
            while True:
                bad = np.abs(w) > 2.0
                if not bad.any():
                    break
                w[bad] = rng.standard_normal(bad.sum())
            theta[off:off + a * b] = w * real
    return theta.astype(dt)


def _forward_net(net, X, dt, act="tanh"):
    acts = [X]
    h = X
    for W, b in net[:-1]:
        h = activation(h @ W.astype(dt) + b.astype(dt), act)
        acts.append(h)
    W, b = net[-1]
    out = (h @ W.astype(dt) + b.astype(dt))[:, 0]
    return out, acts


def forward(theta, dims, X, phase_p, dt=np.float64, act="tanh"):
    """u at positions X (m,3); phase_p (m,) True selects the + network."""
    nets = unflatten(np.asarray(theta, dtype=dt), dims)
    X = np.asarray(X, dtype=dt)
    u = np.zeros(X.shape[0], dtype=dt)
    for tag, sel in (("p", phase_p), ("m", ~phase_p)):
        if sel.any():
            u[sel], _ = _forward_net(nets[tag], X[sel], dt, act)
    return u


def vjp(theta, dims, X, phase_p, seed, dt=np.float64, act="tanh"):
    """Gradient of sum_i seed_i * u(X_i) with respect to the flat parameters."""
    nets = unflatten(np.asarray(theta, dtype=dt), dims)
    lay = layout(dims)
    grad = np.zeros(lay["size"], dtype=dt)
    X = np.asarray(X, dtype=dt)
    seed = np.asarray(seed, dtype=dt)
    for tag, sel in (("p", phase_p), ("m", ~phase_p)):
        if not sel.any():
            continue
        net = nets[tag]
        _, acts = _forward_net(net, X[sel], dt, act)
        d = seed[sel][:, None]  # d out, (m,1)
        for li in range(len(net) - 1, -1, -1):
            W, b = net[li]
            a_in = acts[li]
            off, fa, fb = lay[tag][li]
            grad[off:off + fa * fb] += (a_in.T @ d).reshape(-1)
            grad[off + fa * fb:off + fa * fb + fb] += d.sum(axis=0)
            if li > 0:
                d = (d @ W.astype(dt).T) * activation_grad(acts[li], act)
    return grad


def to_haiku(theta, dims):
    """Flat vector -> the reference's parameter tree (haiku names, MLP.py:101-139)."""
    nets = unflatten(np.asarray(theta), dims)
    tree = {}
    for tag in ("p", "m"):
        for i, (W, b) in enumerate(nets[tag]):
            name = f"double_mlp/~mlp_{tag}_fn/linear" + ("" if i == 0 else f"_{i}")
            tree[name] = {"w": np.array(W), "b": np.array(b)}
    return tree
