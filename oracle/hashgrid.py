"""Multiresolution hash network (hashgrid encoder + CELU coordinate MLPs) —
CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates HashGridEncoder.__call__ (jax_dips/nn/hash_encoding/blocks/
encoders_flax.py:67-254) and SolutionMLP / CoordinateBasedMLP
(nn/hash_encoding/model.py:53-94, blocks/nerfs_flax.py:142-170):

* pos = (x + bound) / (2 bound); level i has scale s_i = N_min * b^i - 1 with
  b = exp((ln N_max - ln N_min) / (L - 1)) (float64 host arithmetic, stored
  float32), resolution res_i = ceil(s_i) + 1; levels with res_i^3 <= T are
  dense (index x + y res + z res^2), the others hashed (x ^ y*2654435761 ^
  z*805459861 in uint32); every index is taken mod T and offset by the level's
  first table row (encoders_flax.py:98-112, :145-197);
* pos_scaled = pos * s_i + 0.5 (float32), corners floor(pos_scaled) + {0,1}^3
  in the order (0,0,0), (0,0,1), ..., (1,1,1); weights prod_d clip((1-o) +
  (2o-1) frac, 0, 1) with frac = modf(pos_scaled) (:199-224);
* encoding[l*F + f] = sum_corners table[idx, f] * weight (:229-235); total
  variation does not enter the solution (model.py:82-94);
* u = MLP_phase(encoding), MLP = [Dense(D), celu] then Dense(1).

Flat parameters: the + MLP, the - MLP (per layer kernel [in][out] row-major,
then bias), then the table (rows x F, row-major).  Table rows are clamped to
the table (jax's default gather clamps out-of-range indices).  Left of the box
(pos_scaled < 0) the reference's corner coordinates go through a float ->
uint32 conversion of negative values (encoders_flax.py:128-132); this
restatement takes that conversion as saturating (negative -> 0, the semantics
of XLA's convert and of the GPU conversion instructions), and modf's negative
fraction puts the whole weight of that axis on the lower corner (parity
unpinned for such positions: the reference's own tests hold no vectors there).

``Frequency`` restates the frequency encoder of the same model type
(encoders_flax.py:269-307), ``SphericalHarmonics`` its spherical-harmonics
encoder (encoders_flax.py:311-538); forward / vjp take any of the encoders.

Pinned by tests/test_oracle_hashgrid.py: dense-level indices and a linear table
reproduce a linear function exactly (trilinear interpolation), a constant
table gives a constant encoding (weights sum to 1), and the VJP against central
finite differences for table rows and MLP weights; the frequency encoding's
closed forms (pos = 0 gives sin 0 = 0 and cos 0 = 1 on every frequency,
pos = 1/2 gives sin(2^l pi / 2)) and its VJP against finite differences; the
spherical harmonics against scipy's complex spherical harmonics on the unit
sphere (ordering, signs, normalisation) and against numpy's Legendre
polynomials off the sphere (the tabulated form, no r^2 terms).
"""
import math

import numpy as np

from . import mlp as M

PRIMES = (1, 2654435761, 805459861)


class HashGrid:
    def __init__(self, L=16, T=2 ** 19, F=2, N_min=16, N_max=2048, bound=1.0):
        self.L, self.T, self.F, self.N_min = int(L), int(T), int(F), int(N_min)
        self.N_max = int(N_max * bound)
        self.bound = float(bound)
        b = math.exp((math.log(self.N_max) - math.log(self.N_min)) / (self.L - 1))
        self.scales, self.res, self.dense, offs = [], [], [], [0]
        for i in range(self.L):
            scale = self.N_min * (b ** i) - 1
            res = math.ceil(scale) + 1
            n = res ** 3
            dense = n <= self.T
            self.scales.append(scale)
            self.res.append(res)
            self.dense.append(dense)
            offs.append(offs[-1] + (n if dense else self.T))
        self.scales32 = np.asarray(self.scales, np.float32)
        self.offsets = np.asarray(offs, np.int64)
        self.rows = int(offs[-1])

    @property
    def n_features(self):
        return self.L * self.F

    def corners(self, X):
        """(L, n, 8) table rows and trilinear weights of positions X (n,3)."""
        X = np.asarray(X, np.float32)
        pos = (X + np.float32(self.bound)) / np.float32(2 * self.bound)
        n = X.shape[0]
        rows = np.zeros((self.L, n, 8), np.int64)
        wts = np.zeros((self.L, n, 8), np.float64)
        for l in range(self.L):
            ps = pos * self.scales32[l] + np.float32(0.5)  # float32
            fl = np.floor(ps)
            frac, _ = np.modf(ps)  # negative below 0 (jnp.modf)
            for c in range(8):
                o = np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1])
                # float -> uint32 conversion of the corner coordinates, saturating below 0
                v = np.maximum(fl + o.astype(np.float32), 0).astype(np.uint64)
                if self.dense[l]:
                    r = np.uint64(self.res[l])
                    idx = (v[:, 0] + v[:, 1] * r + v[:, 2] * r * r) & np.uint64(0xFFFFFFFF)
                else:
                    idx = (v[:, 0] ^ ((v[:, 1] * np.uint64(PRIMES[1])) & np.uint64(0xFFFFFFFF))
                           ^ ((v[:, 2] * np.uint64(PRIMES[2])) & np.uint64(0xFFFFFFFF)))
                idx = (idx % np.uint64(self.T)).astype(np.int64) + self.offsets[l]
                rows[l, :, c] = np.minimum(idx, self.rows - 1)
                w = np.clip((1 - o)[None, :] + (2 * o - 1)[None, :] * frac.astype(np.float64), 0, 1)
                wts[l, :, c] = w.prod(axis=1)
        return rows, wts

    def encode(self, table, X, dt=np.float64):
        rows, wts = self.corners(X)
        tab = np.asarray(table, dt).reshape(self.rows, self.F)
        enc = (tab[rows] * wts[..., None]).sum(axis=2)  # (L, n, F)
        return enc.transpose(1, 0, 2).reshape(X.shape[0], self.L * self.F), (rows, wts)


class Frequency:
    """FrequencyEncoder (encoders_flax.py:269-307) with L frequencies (the hash
    network uses L = 10, model.py:122-123): pos = (x + bound) / (2 bound),
    A[d][l] = (2^l pi) pos_d in float32 (2^l pi is exact), features
    (d L + l) 2 = sin A and (d L + l) 2 + 1 = cos A.  No table (rows = 0)."""

    def __init__(self, L=10, bound=1.0):
        self.L, self.bound = int(L), float(bound)
        self.rows, self.F = 0, 2

    @property
    def n_features(self):
        return 6 * self.L

    def encode(self, table, X, dt=np.float64):
        X = np.asarray(X, np.float32)
        pos = (X + np.float32(self.bound)) / np.float32(2 * self.bound)
        c = np.exp2(np.arange(self.L)).astype(np.float32) * np.float32(np.pi)
        A = (c[None, None, :] * pos[:, :, None]).astype(np.float32)  # (n, 3, L)
        enc = np.stack([np.sin(A.astype(dt)), np.cos(A.astype(dt))], axis=-1)
        return enc.reshape(X.shape[0], 6 * self.L), None


class SphericalHarmonics:
    """SphericalHarmonicsEncoder (encoders_flax.py:311-538) of highest order L
    (1..8, L^2 features), applied to the raw position x (model.py:84 passes x
    itself; the bound is not used).  Feature l^2 + l + m (|m| <= l < L) is

        (-1)^k c_l^k Q_l^k(z) T_m(x, y),   k = |m|,
        c_l^k = sqrt((2l+1)/(4 pi) (l-k)!/(l+k)!) * (sqrt 2 if k > 0),
        Q_l^k = d^k P_l / dz^k (P_l the Legendre polynomial),
        T_m = Re (x + i y)^k for m > 0, Im (x + i y)^k for m < 0, 1 for m = 0,

    the real spherical harmonics in the form the reference tabulates them (the
    unit-vector forms, with no r^2 terms): on the unit sphere they equal
    sqrt 2 Re / Im of the Condon-Shortley Y_l^k.  Q_l^k from the three-term
    recurrence (l - k) Q_l^k = (2l - 1) z Q_{l-1}^k - (l + k - 1) Q_{l-2}^k with
    Q_k^k = (2k - 1)!!, Q_{k+1}^k = (2k + 1) z Q_k^k; (x + iy)^k by complex
    multiplication.  No table (rows = 0)."""

    def __init__(self, L=8, bound=1.0, exact_input=False):
        if not 1 <= int(L) <= 8:
            raise ValueError("spherical harmonics of order 1..8")
        self.L, self.bound = int(L), float(bound)
        self.rows, self.F = 0, 2
        self.exact_input = bool(exact_input)  # positions not rounded to float32 first (finite differences)

    @property
    def n_features(self):
        return self.L * self.L

    @staticmethod
    def coeff(l, k):
        c = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - k) / math.factorial(l + k))
        return (-1) ** k * c * (math.sqrt(2.0) if k > 0 else 1.0)

    def legendre_derivs(self, z):
        """Q[l][k] = d^k P_l / dz^k at z, l, k < L (arrays of z's shape)."""
        L = self.L
        Q = [[np.zeros_like(z) for _ in range(L)] for _ in range(L)]
        for k in range(L):
            dfact = 1.0
            for j in range(1, 2 * k, 2):
                dfact *= j
            Q[k][k] = np.full_like(z, dfact)
            if k + 1 < L:
                Q[k + 1][k] = (2 * k + 1) * z * Q[k][k]
            for l in range(k + 2, L):
                Q[l][k] = ((2 * l - 1) * z * Q[l - 1][k] - (l + k - 1) * Q[l - 2][k]) / (l - k)
        return Q

    def encode(self, table, X, dt=np.float64):
        X = np.asarray(X, np.float64 if self.exact_input else np.float32).astype(dt)
        x, y, z = X[:, 0], X[:, 1], X[:, 2]
        Q = self.legendre_derivs(z)
        C, S = [np.ones_like(x)], [np.zeros_like(x)]
        for k in range(1, self.L):
            C.append(x * C[-1] - y * S[-1])
            S.append(x * S[-1] + y * C[-2])
        enc = np.zeros((X.shape[0], self.L * self.L), dt)
        for l in range(self.L):
            for m in range(-l, l + 1):
                k = abs(m)
                t = C[k] if m > 0 else (S[k] if m < 0 else 1.0)
                enc[:, l * l + l + m] = self.coeff(l, k) * Q[l][k] * t
        return enc, None


def mlp_size(E, widths):
    dims = [E] + list(widths) + [1]
    return sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))


def _unflatten_mlp(flat, E, widths):
    dims = [E] + list(widths) + [1]
    out, off = [], 0
    for i in range(len(dims) - 1):
        a, b = dims[i], dims[i + 1]
        out.append((np.asarray(flat[off:off + a * b]).reshape(a, b), np.asarray(flat[off + a * b:off + a * b + b])))
        off += a * b + b
    return out


def split(theta, grid, widths):
    E = grid.n_features
    P = mlp_size(E, widths)
    return theta[:P], theta[P:2 * P], theta[2 * P:]


def forward(theta, grid, widths, X, phase_p, dt=np.float64):
    """u at X (n,3); phase_p True selects the + network."""
    tp, tm, tab = split(np.asarray(theta, dt), grid, widths)
    enc, _ = grid.encode(tab, X, dt)
    u = np.zeros(X.shape[0], dt)
    for flat, sel in ((tp, phase_p), (tm, ~phase_p)):
        if sel.any():
            u[sel], _ = M._forward_net(_unflatten_mlp(flat, grid.n_features, widths), enc[sel], dt, "celu")
    return u


def vjp(theta, grid, widths, X, phase_p, seed, dt=np.float64):
    """Gradient of sum_i seed_i u(X_i) with respect to the flat parameters."""
    theta = np.asarray(theta, dt)
    E = grid.n_features
    P = mlp_size(E, widths)
    tp, tm, tab = split(theta, grid, widths)
    enc, aux = grid.encode(tab, X, dt)
    grad = np.zeros(theta.size, dt)
    denc = np.zeros_like(enc)
    seed = np.asarray(seed, dt)
    for k, (flat, sel) in enumerate(((tp, phase_p), (tm, ~phase_p))):
        if not sel.any():
            continue
        net = _unflatten_mlp(flat, E, widths)
        _, acts = M._forward_net(net, enc[sel], dt, "celu")
        d = seed[sel][:, None]
        g = np.zeros(P, dt)
        offs = []
        off = 0
        for W, b in net:
            offs.append(off)
            off += W.size + b.size
        for li in range(len(net) - 1, -1, -1):
            W, b = net[li]
            g[offs[li]:offs[li] + W.size] += (acts[li].T @ d).reshape(-1)
            g[offs[li] + W.size:offs[li] + W.size + b.size] += d.sum(axis=0)
            d = d @ W.T
            if li > 0:
                d = d * M.activation_grad(acts[li], "celu")
        denc[sel] = d
        grad[k * P:(k + 1) * P] = g
    if not grid.rows:
        return grad
    rows, wts = aux
    # table: encoding (n, L, F) <- sum_c table[rows[l, :, c]] * wts[l, :, c]
    gt = np.zeros((grid.rows, grid.F), dt)
    de = denc.reshape(X.shape[0], grid.L, grid.F)
    for l in range(grid.L):
        for c in range(8):
            np.add.at(gt, rows[l, :, c], de[:, l, :] * wts[l, :, c][:, None])
    grad[2 * P:] = gt.reshape(-1)
    return grad


def init_params(grid, widths, seed=0, dt=np.float32):
    """Glorot-uniform MLP kernels, zero biases, table U(-1e-4, 1e-4)
    (encoders_flax.py:118-123; none for the frequency encoder); numpy's
    generator."""
    rng = np.random.default_rng(seed)
    parts = []
    dims = [grid.n_features] + list(widths) + [1]
    for _ in range(2):
        for i in range(len(dims) - 1):
            a, b = dims[i], dims[i + 1]
            lim = np.sqrt(6.0 / (a + b))
            parts += [rng.uniform(-lim, lim, a * b), np.zeros(b)]
    if grid.rows:
        parts.append(rng.uniform(-1e-4, 1e-4, grid.rows * grid.F))
    return np.concatenate(parts).astype(dt)


def loss_and_grad(theta, grid, widths, rec, rows, dt=np.float64):
    """0.5 * mean(r^2) over the batch rows of oracle.trainer.LevelRecords rec and
    its gradient, with the hashgrid network (the trainer's loss,
    trainer.py:892-912)."""
    W, nz, X, ph = rec.entries(rows)
    u = forward(theta, grid, widths, X, ph, dt)
    r = rec.c[rows].astype(dt).copy()
    np.add.at(r, nz[0], W[nz].astype(dt) * u)
    B = len(rows)
    loss = np.sum(dt(0.5) * r * r) / dt(B)
    seed = r[nz[0]] * W[nz].astype(dt) / dt(B)
    return loss, vjp(theta, grid, widths, X, ph, seed, dt), r


class Network:
    """forward / vjp of the hashgrid network with a fixed encoder and widths, in
    the call form of oracle.refpattern.step's ``net``."""

    def __init__(self, grid, widths):
        self.grid, self.widths = grid, list(widths)

    def forward(self, theta, X, phase_p, dt=np.float64):
        return forward(theta, self.grid, self.widths, X, phase_p, dt)

    def vjp(self, theta, X, phase_p, seed, dt=np.float64):
        return vjp(theta, self.grid, self.widths, X, phase_p, seed, dt)
