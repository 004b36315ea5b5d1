"""Interface problems of the benchmark workloads, as user callbacks.

Every callback takes coordinate-first positions ``r`` (shape (3, m)) and uses
the ``xnp`` namespace, so the same function samples on numpy (host) or torch
(device) arrays.  The manufactured pairs follow the recipe of the reference's
test configurations (tests/confs/experiment_configs.py:21-412): f = k u - div(mu
grad u) on each side, alpha = u+ - u-, and beta in the reference's sign
convention beta = -(mu+ grad u+ - mu- grad u-) . n (experiment_configs.py:85-96,
the explicit ``* (-1.0)``), with n the normalised level-set gradient and a
nan_to_num guard where that gradient vanishes (experiment_configs.py:295).
phi is perturbed by sign+-(phi) * 1e-10 (geometry/level_set.py:34-48).
"""
import numpy as np

from . import xnp


def perturb(phi):
    """phi + sign+-(phi) * 1e-10 with sign+-(0) = -1 (level_set.py:34-48)."""

    def f(r):
        v = phi(r)
        return v + xnp.where(v > 0, 1.0, -1.0) * 1e-10

    return f


def _u_pair():
    """u- = exp(-|x|^2), u+ = sin x cos y sin z and their gradients/laplacians."""

    def um(r):
        return xnp.exp(-(r[0] ** 2 + r[1] ** 2 + r[2] ** 2))

    def up(r):
        return xnp.sin(r[0]) * xnp.cos(r[1]) * xnp.sin(r[2])

    def grad_um(r):
        e = um(r)
        return (-2 * r[0] * e, -2 * r[1] * e, -2 * r[2] * e)

    def grad_up(r):
        return (xnp.cos(r[0]) * xnp.cos(r[1]) * xnp.sin(r[2]),
                -xnp.sin(r[0]) * xnp.sin(r[1]) * xnp.sin(r[2]),
                xnp.sin(r[0]) * xnp.cos(r[1]) * xnp.cos(r[2]))

    def lap_um(r):
        q = r[0] ** 2 + r[1] ** 2 + r[2] ** 2
        return (4 * q - 6) * um(r)

    def lap_up(r):
        return -3 * up(r)

    return um, up, grad_um, grad_up, lap_um, lap_up


def _const(v):
    return lambda r: xnp.zeros_like(r[0]) + v


def make_problem(phi_raw, grad_phi, mu_m, mu_p, k_m, k_p, name):
    """Assemble the 11 callbacks for constant mu/k and the standard u pair."""
    um, up, gum, gup, lum, lup = _u_pair()
    phi = perturb(phi_raw)

    def f_m(r):
        return k_m * um(r) - mu_m * lum(r)

    def f_p(r):
        return k_p * up(r) - mu_p * lup(r)

    def alpha(r):
        return up(r) - um(r)

    def beta(r):
        g = grad_phi(r)
        nn = xnp.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
        a, b = gup(r), gum(r)
        v = sum((mu_p * a[i] - mu_m * b[i]) * (g[i] / nn) for i in range(3))
        return xnp.nan_to_num(-v)

    def exact(r):
        return xnp.where(phi(r) >= 0, up(r), um(r))

    return dict(
        name=name,
        initial_value_fn=_const(0.0),
        dirichlet_bc_fn=exact,
        lvl_set_fn=phi,
        mu_m_fn=_const(mu_m),
        mu_p_fn=_const(mu_p),
        k_m_fn=_const(k_m),
        k_p_fn=_const(k_p),
        f_m_fn=f_m,
        f_p_fn=f_p,
        alpha_fn=alpha,
        beta_fn=beta,
        exact_fn=exact,
        exact_m_fn=um,
        exact_p_fn=up,
    )


def sphere(radius=0.5, mu_m=1.0, mu_p=10.0, k_m=0.0, k_p=0.0, centre=(0.0, 0.0, 0.0)):
    """phi = |x - c| - radius (the 16^3 CPU, 64^3 and 128^3 sphere workloads)."""
    cx, cy, cz = centre

    def phi(r):
        return xnp.sqrt((r[0] - cx) ** 2 + (r[1] - cy) ** 2 + (r[2] - cz) ** 2) - radius

    def gphi(r):
        return (r[0] - cx, r[1] - cy, r[2] - cz)

    return make_problem(phi, gphi, mu_m, mu_p, k_m, k_p, "sphere")


def star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0, centre=(0.0, 0.0, 0.0)):
    """phi = r - (r0 + amp sin(lobes_t theta) sin(lobes_p varphi)) about ``centre``,
    theta = arccos(z/r), varphi = atan2(y, x); r = 0 guarded (phi = -r0)."""
    cx, cy, cz = centre

    def phi(r):
        x, y, z = r[0] - cx, r[1] - cy, r[2] - cz
        rr = xnp.sqrt(x ** 2 + y ** 2 + z ** 2)
        th = xnp.arccos(xnp.nan_to_num(z / rr))
        ph = xnp.arctan2(y, x)
        return rr - (r0 + amp * xnp.sin(lobes_t * th) * xnp.sin(lobes_p * ph))

    def gphi(r):
        x, y, z = r[0] - cx, r[1] - cy, r[2] - cz
        rr = xnp.sqrt(x ** 2 + y ** 2 + z ** 2)
        rho = xnp.sqrt(x ** 2 + y ** 2)
        th = xnp.arccos(xnp.nan_to_num(z / rr))
        ph = xnp.arctan2(y, x)
        dth = amp * lobes_t * xnp.cos(lobes_t * th) * xnp.sin(lobes_p * ph)
        dph = amp * lobes_p * xnp.sin(lobes_t * th) * xnp.cos(lobes_p * ph)
        # grad theta = (xz, yz, -rho^2) / (r^2 rho); grad varphi = (-y, x, 0) / rho^2
        gx = x / rr - dth * (x * z) / (rr ** 2 * rho) - dph * (-y) / rho ** 2
        gy = y / rr - dth * (y * z) / (rr ** 2 * rho) - dph * x / rho ** 2
        gz = z / rr - dth * (-(rho ** 2)) / (rr ** 2 * rho)
        return (gx, gy, gz)

    return make_problem(phi, gphi, mu_m, mu_p, k_m, k_p, "star")


class AtomField:
    """Union of spheres and a sum of Gaussian charges, evaluated through a cell
    list so that a query reads only the atoms of its 27 neighbouring cells.

    phi(x) = min(cap, min_i |x - c_i| - r_i) over the atoms of the 27 cells
    around x: with cell size >= cap + r_max every atom outside them is at
    least cap away from the surface term, so the value equals the global
    minimum wherever it is below cap (the geometry and regression read phi
    values only within a few grid spacings of the interface; elsewhere only the
    sign matters).  charge(x) = sum_i q_i N(x; c_i, sigma^2 I) over the same
    atoms (the Gaussians are < 1e-13 of their peak beyond one cell).
    Works on numpy arrays and torch tensors of shape (3, m)."""

    def __init__(self, centres, radii, charges, lo, hi, cap=2.0, sigma=0.5, chunk=1 << 20):
        self.c = np.asarray(centres, np.float64)
        self.r = np.asarray(radii, np.float64)
        self.q = np.asarray(charges, np.float64)
        self.cap, self.sigma, self.chunk = float(cap), float(sigma), int(chunk)
        cell = max(self.cap + float(self.r.max()), 4.0 * self.sigma)
        self.lo = float(lo)
        self.nc = max(1, int(np.ceil((float(hi) - float(lo)) / cell)))
        self.cell = (float(hi) - float(lo)) / self.nc
        if self.cell < self.cap + float(self.r.max()):  # rounding of nc
            self.nc -= 1
            self.cell = (float(hi) - float(lo)) / max(self.nc, 1)
        idx = np.clip(((self.c - self.lo) / self.cell).astype(np.int64), 0, self.nc - 1)
        flat = (idx[:, 0] * self.nc + idx[:, 1]) * self.nc + idx[:, 2]
        counts = np.bincount(flat, minlength=self.nc ** 3)
        K = max(1, int(counts.max()))
        # padded per-cell atom lists; padding atoms sit far away with zero charge
        table = np.full((self.nc ** 3, K), len(self.c), np.int64)
        fill = np.zeros(self.nc ** 3, np.int64)
        for a, f in enumerate(flat):
            table[f, fill[f]] = a
            fill[f] += 1
        self.table = table
        self.cx = np.concatenate([self.c, [[1e6, 1e6, 1e6]]])
        self.rx = np.concatenate([self.r, [0.0]])
        self.qx = np.concatenate([self.q, [0.0]])
        self._dev = {}

    def _arrays(self, like):
        if not xnp._is_torch(like):
            return np, self.table, self.cx, self.rx, self.qx
        import torch

        key = str(like.device)
        if key not in self._dev:
            self._dev[key] = tuple(torch.as_tensor(a, device=like.device) for a in
                                   (self.table, self.cx, self.rx, self.qx))
        return (torch,) + self._dev[key]

    def _eval(self, r, what):
        lib, table, cx, rx, qx = self._arrays(r)
        m = r.shape[1]
        out = lib.empty(m, dtype=lib.float64) if lib is np else lib.empty(m, dtype=lib.float64, device=r.device)
        offs = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)], np.int64)
        if lib is not np:
            offs = lib.as_tensor(offs, device=r.device)
        for a in range(0, m, self.chunk):
            b = min(m, a + self.chunk)
            x = r[:, a:b].T  # (n,3)
            ci = ((x - self.lo) / self.cell)
            ci = ci.astype(np.int64) if lib is np else ci.floor().to(lib.int64)
            ci = ci.clip(0, self.nc - 1) if lib is np else ci.clamp(0, self.nc - 1)
            nb = ci[:, None, :] + offs[None]  # (n,27,3)
            ok = ((nb >= 0) & (nb < self.nc)).all(axis=2) if lib is np else ((nb >= 0) & (nb < self.nc)).all(dim=2)
            nbc = nb.clip(0, self.nc - 1) if lib is np else nb.clamp(0, self.nc - 1)
            cellid = (nbc[..., 0] * self.nc + nbc[..., 1]) * self.nc + nbc[..., 2]
            atoms = table[cellid]  # (n,27,K)
            pad = len(self.c)
            atoms = lib.where(ok[..., None], atoms, lib.full_like(atoms, pad) if lib is not np else pad)
            atoms = atoms.reshape(atoms.shape[0], -1)
            d = cx[atoms] - x[:, None, :]
            dist = (d * d).sum(axis=2) if lib is np else (d * d).sum(dim=2)
            if what == "phi":
                v = lib.sqrt(dist) - rx[atoms]
                v = v.min(axis=1) if lib is np else v.min(dim=1).values
                v = lib.minimum(v, lib.full_like(v, self.cap)) if lib is not np else np.minimum(v, self.cap)
            else:
                s2 = self.sigma ** 2
                g = qx[atoms] * lib.exp(-dist / (2.0 * s2)) / (2.0 * np.pi * s2) ** 1.5
                v = g.sum(axis=1) if lib is np else g.sum(dim=1)
            out[a:b] = v
        return out

    def phi(self, r):
        return self._eval(r, "phi")

    def charge(self, r):
        return self._eval(r, "charge")


def pb_box(n_atoms=2000, half_width=20.0, seed=0, mu_m=2.0, mu_p=80.0, kappa2=0.015625, r_range=(1.2, 2.0)):
    """Poisson-Boltzmann-like box problem (Angstrom units): n_atoms spheres with
    centres uniform in [-w, w]^3 and radii uniform in r_range
    (numpy default_rng(seed)); phi = min_i |x - c_i| - r_i (capped, see
    AtomField), mu- = 2 inside, mu+ = 80 outside, k- = 0, k+ = kappa^2,
    f- = sum_i q_i N(x; c_i, 0.5^2 I) with q_i = +-1, f+ = 0, alpha = beta = 0,
    g = 0.  No analytic solution (the grid judge is the reference answer)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-half_width, half_width, (n_atoms, 3))
    rad = rng.uniform(r_range[0], r_range[1], n_atoms)
    q = rng.choice((-1.0, 1.0), n_atoms)
    field = AtomField(c, rad, q, -half_width, half_width)
    zero = _const(0.0)
    return dict(
        name="pb",
        initial_value_fn=zero,
        dirichlet_bc_fn=zero,
        lvl_set_fn=perturb(field.phi),
        mu_m_fn=_const(mu_m),
        mu_p_fn=_const(mu_p),
        k_m_fn=_const(0.0),
        k_p_fn=_const(kappa2),
        f_m_fn=field.charge,
        f_p_fn=zero,
        alpha_fn=zero,
        beta_fn=zero,
        field=field,
    )


SWEEP_SIZES = (32, 48, 64, 96)


def sweep_problem(i):
    """Problem i of the 64-problem sweep workload: numpy default_rng(seed=i)
    draws N from {32^3, 48^3, 64^3, 96^3}, a sphere (radius in [0.3, 0.7]) or a
    star (amplitude in [0.05, 0.15], lobes in {3, 5, 7}), mu+ log-uniform in
    [2, 1000], mu- = 1, k = 0, and the manufactured solution pair of the sphere
    workloads.  Returns (problem, n)."""
    rng = np.random.default_rng(seed=i)
    n = int(rng.choice(SWEEP_SIZES))
    mu_p = float(np.exp(rng.uniform(np.log(2.0), np.log(1000.0))))
    if rng.uniform() < 0.5:
        prob = sphere(radius=float(rng.uniform(0.3, 0.7)), mu_m=1.0, mu_p=mu_p, k_m=0.0, k_p=0.0)
    else:
        lobes = int(rng.choice((3, 5, 7)))
        prob = star(r0=0.5, amp=float(rng.uniform(0.05, 0.15)), lobes_t=lobes, lobes_p=lobes, mu_m=1.0, mu_p=mu_p,
                    k_m=0.0, k_p=0.0)
    return prob, n


def grid_axes(n, lo=-1.0, hi=1.0):
    """float32 linspace as the reference builds grids (README quick example)."""
    a = np.linspace(lo, hi, n, dtype=np.float32)
    return a, a.copy(), a.copy()


def oracle_callbacks(prob):
    """The callback dict in the name scheme of the CPU oracle."""
    return dict(phi=prob["lvl_set_fn"], mu_m=prob["mu_m_fn"], mu_p=prob["mu_p_fn"],
                k_m=prob["k_m_fn"], k_p=prob["k_p_fn"], f_m=prob["f_m_fn"], f_p=prob["f_p_fn"],
                alpha=prob["alpha_fn"], beta=prob["beta_fn"], g=prob["dirichlet_bc_fn"])
