"""Pins the oracle's network (oracle/mlp.py), optimizer (oracle/optim.py) and
schedule/batching (oracle/trainer.py):
* forward against an independent torch (float64) network; the hand VJP against
  torch autograd and against central finite differences (h = 1e-3, 32 params);
* Adam (no clipping) against torch.optim.Adam on the same gradient sequence;
  clip_by_global_norm and the exponential / polynomial schedules against their
  closed forms;
* zoom levels and contiguous batches.
"""
import numpy as np
import pytest
import torch

from oracle import mlp as M
from oracle import optim as O
from oracle import trainer as T


def torch_net(theta, dims):
    nets = M.unflatten(torch.tensor(theta, dtype=torch.float64), dims)

    def f(X, phase):
        out = torch.zeros(X.shape[0], dtype=torch.float64)
        for tag, sel in (("p", phase), ("m", ~phase)):
            h = X[sel]
            for W, b in nets[tag][:-1]:
                h = torch.tanh(h @ W + b)
            W, b = nets[tag][-1]
            out[sel] = (h @ W + b)[:, 0]
        return out

    return f, nets


@pytest.mark.parametrize("act", ["tanh", "celu"])
@pytest.mark.parametrize("dims", [(1, 1, 2, 10), (2, 16, 2, 16), (3, 8, 1, 5)])
def test_forward_and_vjp(dims, act):
    rng = np.random.default_rng(0)
    theta = M.init_params(dims, seed=4).astype(np.float64)
    X = rng.uniform(-1, 1, size=(300, 3))
    phase = rng.uniform(size=300) > 0.4
    seed = rng.normal(size=300)
    u = M.forward(theta, dims, X, phase, act=act)
    th = torch.tensor(theta, requires_grad=True)
    tact = torch.tanh if act == "tanh" else torch.nn.functional.celu
    nets = M.unflatten(th, dims)
    Xt = torch.tensor(X)
    pt = torch.tensor(phase)
    out = torch.zeros(300, dtype=torch.float64)
    for tag, sel in (("p", pt), ("m", ~pt)):
        h = Xt[sel]
        for W, b in nets[tag][:-1]:
            h = tact(h @ W + b)
        W, b = nets[tag][-1]
        out = out.index_put((torch.nonzero(sel)[:, 0],), (h @ W + b)[:, 0])
    assert np.allclose(u, out.detach().numpy(), rtol=1e-13, atol=1e-13)
    (out * torch.tensor(seed)).sum().backward()
    g = M.vjp(theta, dims, X, phase, seed, act=act)
    assert np.allclose(g, th.grad.numpy(), rtol=1e-11, atol=1e-12)
    # CELU's second derivative jumps at 0: a smaller step keeps the central
    # difference's O(h) error near the kink below the bar
    h = 1e-3 if act == "tanh" else 1e-5
    for i in rng.choice(theta.size, 32, replace=False):
        e = np.zeros_like(theta)
        e[i] = h
        fd = (seed @ M.forward(theta + e, dims, X, phase, act=act)
              - seed @ M.forward(theta - e, dims, X, phase, act=act)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i]))


def test_adam_matches_torch():
    rng = np.random.default_rng(1)
    p0 = rng.normal(size=50).astype(np.float32)
    grads = [rng.normal(size=50).astype(np.float32) * 0.1 for _ in range(30)]
    cfg = {"optimizer_name": "adam", "learning_rate": 1e-2}
    st = O.OptState(50, np.float32)
    p = p0.copy()
    for g in grads:
        p = O.step(p, g, st, cfg, np.float32)
    tp = torch.nn.Parameter(torch.tensor(p0))
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
    assert np.allclose(p, tp.detach().numpy(), rtol=2e-5, atol=2e-7)


def test_clip_and_schedules():
    g = np.array([3.0, 4.0], np.float32)  # norm 5 -> clipped to 1
    st = O.OptState(2, np.float32)
    cfg = {"optimizer_name": "custom", "learning_rate": 0.5, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    p = O.step(np.zeros(2, np.float32), g, st, cfg, np.float32)
    # first Adam step on the clipped gradient gives -lr * sign(g)
    assert np.allclose(p, -0.5 * np.sign(g), rtol=1e-5)
    assert np.isclose(O.schedule("exponential", 1e-3, 0.9, 0, np.float64), 1e-3)
    assert np.isclose(O.schedule("exponential", 1e-3, 0.9, 1500, np.float64), 1e-3 * 0.9 ** 1.5)
    assert np.isclose(O.schedule("polynomial", 1e-3, 0.9, 250, np.float64), 0.75e-3)
    assert np.isclose(O.schedule("polynomial", 1e-3, 0.9, 5000, np.float64), 0.0)


def test_zoom_levels_and_batches():
    assert [T.zoom_level(e, 8) for e in range(8)] == [0, 0, 1, 1, 2, 2, 3, 3]
    assert [T.zoom_level(e, 10) for e in range(10)] == [0, 0, 1, 1, 2, 2, 3, 3, 4, 4]
    with pytest.raises(ZeroDivisionError):
        T.zoom_level(0, 3)
    h = T.level_spacing(np.array([0.25, 0.5, 0.125], np.float32), 3)
    assert np.array_equal(h, np.array([0.25, 0.5, 0.125], np.float32) / 8)
    b = T.batches(10, 4)
    assert [list(x) for x in b] == [[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, -1, -1]]


def test_celu_closed_form():
    """CELU (alpha 1): identity above 0, expm1 below, continuous with slope 1 at 0."""
    x = np.array([-30.0, -2.0, -1e-8, 0.0, 1e-8, 3.0])
    a = M.activation(x, "celu")
    assert np.array_equal(a[3:], x[3:]) and np.allclose(a[:3], np.expm1(x[:3]), rtol=1e-15)
    assert np.allclose(M.activation_grad(a, "celu"), np.where(x > 0, 1.0, np.exp(x)), rtol=1e-12)
