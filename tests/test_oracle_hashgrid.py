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


def _sh_scipy(L, X):
    """Real spherical harmonics of unit vectors X from scipy's complex ones
    (Condon-Shortley phase): sqrt 2 Re Y_l^k for m = k > 0, sqrt 2 Im Y_l^k for
    m = -k < 0, Y_l^0 for m = 0, feature l^2 + l + m."""
    from scipy.special import sph_harm_y

    polar = np.arccos(np.clip(X[:, 2], -1, 1))
    azim = np.arctan2(X[:, 1], X[:, 0])
    out = np.zeros((X.shape[0], L * L))
    for l in range(L):
        for m in range(-l, l + 1):
            y = sph_harm_y(l, abs(m), polar, azim)
            out[:, l * l + l + m] = (np.sqrt(2) * (y.real if m > 0 else y.imag)) if m else y.real
    return out


def test_spherical_harmonics_on_the_sphere_and_off_it():
    """On unit vectors the encoding equals scipy's real spherical harmonics
    (ordering, signs, normalisation; the reference's first terms are
    1/(2 sqrt pi), -sqrt3 y/(2 sqrt pi), sqrt3 z/(2 sqrt pi), -sqrt3 x/(2 sqrt pi));
    off the sphere it is c Q(z) T(x, y) with Q the derivatives of numpy's
    Legendre polynomials and T from numpy's complex powers -- the unit-vector
    form the reference evaluates on raw positions, without r^2 terms (the solid
    harmonic 3 z^2 - r^2 differs from 3 z^2 - 1 there)."""
    from numpy.polynomial import legendre

    rng = np.random.default_rng(9)
    U = rng.normal(size=(300, 3))
    U /= np.linalg.norm(U, axis=1, keepdims=True)
    sh = HG.SphericalHarmonics(8)
    assert sh.n_features == 64
    enc, _ = sh.encode(None, U.astype(np.float32))
    Uf = U.astype(np.float32).astype(np.float64)
    Uf /= np.linalg.norm(Uf, axis=1, keepdims=True)
    assert np.max(np.abs(enc - _sh_scipy(8, Uf))) <= 1e-5
    s = 1.0 / (2.0 * np.sqrt(np.pi))
    assert np.allclose(enc[:, 0], s)
    assert np.allclose(enc[:, 1:4], np.sqrt(3) * s * np.stack([-Uf[:, 1], Uf[:, 2], -Uf[:, 0]], 1), atol=1e-6)
    X = rng.uniform(-1.6, 1.6, (200, 3)).astype(np.float32)
    enc, _ = sh.encode(None, X)
    Xd = X.astype(np.float64)
    w = (Xd[:, 0] + 1j * Xd[:, 1])
    for l in range(8):
        for m in range(-l, l + 1):
            k = abs(m)
            q = legendre.Legendre.basis(l).deriv(k)(Xd[:, 2]) if k <= l else 0.0
            t = (w ** k).real if m > 0 else ((w ** k).imag if m < 0 else 1.0)
            want = HG.SphericalHarmonics.coeff(l, k) * q * t
            assert np.allclose(enc[:, l * l + l + m], want, rtol=1e-9, atol=1e-9), (l, m)
    # lower orders are the leading features of the higher ones
    e2, _ = HG.SphericalHarmonics(2).encode(None, X)
    assert np.array_equal(e2, enc[:, :4])
    with pytest.raises(ValueError):
        HG.SphericalHarmonics(9)


def test_spherical_harmonics_network_vjp():
    sh = HG.SphericalHarmonics(3)
    widths = [16]
    theta = HG.init_params(sh, widths, seed=4).astype(np.float64)
    assert theta.size == 2 * HG.mlp_size(9, widths)
    rng = np.random.default_rng(7)
    X = rng.uniform(-0.9, 0.9, (40, 3)).astype(np.float32)
    phase = rng.uniform(size=40) > 0.5
    seed = rng.normal(size=40)
    gr = HG.vjp(theta, sh, widths, X, phase, seed)
    h = 1e-5
    for i in rng.choice(theta.size, 24, replace=False):
        e = np.zeros_like(theta)
        e[i] = h
        fd = (seed @ HG.forward(theta + e, sh, widths, X, phase) - seed @ HG.forward(theta - e, sh, widths, X, phase)) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-5 * max(1.0, abs(gr[i])), (i, fd, gr[i])
