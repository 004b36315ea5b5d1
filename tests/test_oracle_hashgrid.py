"""Pins the oracle's hashgrid network (oracle/hashgrid.py) without the
reference: closed forms of the encoder and finite differences of the VJP."""
import numpy as np
import pytest

from oracle import hashgrid as HG


def _grid():
    return HG.HashGrid(L=4, T=2 ** 10, F=2, N_min=4, N_max=32, bound=1.0)


def test_levels_and_offsets():
    g = HG.HashGrid()  # the reference defaults: L 16, T 2^19, N_min 16, N_max 2048
    assert g.res[0] == 16 and g.dense[0] and not g.dense[-1]
    assert all(a or not b for a, b in zip(g.dense[:-1], g.dense[1:]))  # dense levels come first
    assert np.all(np.diff(g.offsets) == [min(r ** 3, g.T) for r in g.res])
    assert np.isclose(g.scales[-1], 2047.0)


def test_constant_and_linear_tables():
    g = _grid()
    rng = np.random.default_rng(0)
    X = rng.uniform(-0.95, 0.95, (500, 3)).astype(np.float32)
    # constant table: the trilinear weights of a position sum to 1
    tab = np.ones((g.rows, 2)) * 0.25
    enc, (rows, wts) = g.encode(tab, X)
    assert np.allclose(enc, 0.25, rtol=1e-12) and np.allclose(wts.sum(axis=2), 1.0)
    # linear function of the vertex coordinates on a dense level: interpolation is exact
    # (positions whose corners stay below res; a corner at coordinate res aliases
    # into the next row of the level, as the reference's index arithmetic does)
    X = rng.uniform(-0.6, 0.6, (500, 3)).astype(np.float32)
    for l in range(g.L):
        if not g.dense[l]:
            continue
        r = g.res[l]
        idx = np.arange(r ** 3)
        vx, vy, vz = idx % r, (idx // r) % r, idx // (r * r)
        tab = np.zeros((g.rows, 2))
        tab[g.offsets[l] + idx, 0] = 2.0 * vx - 3.0 * vy + 0.5 * vz
        enc, _ = g.encode(tab, X)
        pos = (X + np.float32(1.0)) / np.float32(2.0)
        ps = (pos * g.scales32[l] + np.float32(0.5)).astype(np.float64)
        want = 2.0 * ps[:, 0] - 3.0 * ps[:, 1] + 0.5 * ps[:, 2]
        assert np.allclose(enc[:, 2 * l], want, rtol=1e-9, atol=1e-9)


def test_vjp_matches_finite_differences():
    g = _grid()
    widths = [16]
    theta = HG.init_params(g, widths, seed=3).astype(np.float64)
    P = HG.mlp_size(g.n_features, widths)
    # a table large enough in magnitude that the MLP input matters
    theta[2 * P:] = np.random.default_rng(4).uniform(-0.5, 0.5, theta.size - 2 * P)
    rng = np.random.default_rng(5)
    X = rng.uniform(-0.9, 0.9, (64, 3)).astype(np.float32)
    phase = rng.uniform(size=64) > 0.5
    seed = rng.normal(size=64)
    gr = HG.vjp(theta, g, widths, X, phase, seed)
    rows, _ = g.corners(X)
    touched = np.unique(rows.reshape(-1))
    idx = np.concatenate([rng.choice(2 * P, 16, replace=False),
                          2 * P + 2 * rng.choice(touched, 12, replace=False),
                          2 * P + 1 + 2 * rng.choice(touched, 12, replace=False)])
    h = 1e-5
    for i in idx:
        e = np.zeros_like(theta)
        e[i] = h
        fd = (seed @ HG.forward(theta + e, g, widths, X, phase) - seed @ HG.forward(theta - e, g, widths, X, phase)) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-5 * max(1.0, abs(gr[i])), (i, fd, gr[i])
    # rows no position touches get no gradient
    untouched = np.setdiff1d(np.arange(g.rows), touched)
    assert np.all(gr[2 * P:].reshape(-1, 2)[untouched] == 0)


def test_left_of_box_saturates_to_the_first_vertex():
    """pos_scaled < 0 on an axis: the float -> uint32 conversion saturates the
    corner coordinates to 0 and modf's negative fraction puts the axis' whole
    weight on the lower corner, so a linear table reads the value at vertex
    coordinate 0 along that axis."""
    g = _grid()
    rng = np.random.default_rng(7)
    X = rng.uniform(-0.6, 0.6, (300, 3)).astype(np.float32)
    X[:, 0] = rng.uniform(-2.6, -2.2, 300).astype(np.float32)  # pos < 0 for every level
    for l in range(g.L):
        if not g.dense[l]:
            continue
        r = g.res[l]
        idx = np.arange(r ** 3)
        vx, vy, vz = idx % r, (idx // r) % r, idx // (r * r)
        tab = np.zeros((g.rows, 2))
        tab[g.offsets[l] + idx, 1] = 2.0 * vx - 3.0 * vy + 0.5 * vz + 1.0
        enc, (rows, wts) = g.encode(tab, X)
        pos = (X + np.float32(1.0)) / np.float32(2.0)
        ps = (pos * g.scales32[l] + np.float32(0.5)).astype(np.float64)
        assert np.all(ps[:, 0] < 0)
        want = -3.0 * ps[:, 1] + 0.5 * ps[:, 2] + 1.0
        assert np.allclose(enc[:, 2 * l + 1], want, rtol=1e-9, atol=1e-9)
        # the upper x corners carry no weight
        assert np.all(wts[l][:, 4:] == 0)


def test_frequency_encoding_closed_forms_and_vjp():
    f = HG.Frequency(10, bound=1.0)
    assert f.n_features == 60
    X = np.array([[-1.0, 0.0, -1.0]], np.float32)  # pos = (0, 1/2, 0)
    enc, _ = f.encode(None, X)
    e = enc.reshape(3, 10, 2)
    assert np.allclose(e[0, :, 0], 0.0) and np.allclose(e[0, :, 1], 1.0)
    want = np.sin(np.exp2(np.arange(10)) * np.float64(np.float32(np.pi)) * 0.5)
    assert np.allclose(e[1, :, 0], want, atol=1e-6)
    widths = [16]
    theta = HG.init_params(f, widths, seed=3).astype(np.float64)
    assert theta.size == 2 * HG.mlp_size(60, widths)
    rng = np.random.default_rng(6)
    X = rng.uniform(-0.9, 0.9, (40, 3)).astype(np.float32)
    phase = rng.uniform(size=40) > 0.5
    seed = rng.normal(size=40)
    gr = HG.vjp(theta, f, widths, X, phase, seed)
    h = 1e-5
    for i in rng.choice(theta.size, 24, replace=False):
        e = np.zeros_like(theta)
        e[i] = h
        fd = (seed @ HG.forward(theta + e, f, widths, X, phase) - seed @ HG.forward(theta - e, f, widths, X, phase)) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-5 * max(1.0, abs(gr[i])), (i, fd, gr[i])
