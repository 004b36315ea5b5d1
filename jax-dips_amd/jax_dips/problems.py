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


def _problem(phi_raw, grad_phi, mu_m, mu_p, grad_mu_m, grad_mu_p, um, up, grad_um, grad_up, f_m, f_p, name,
             normalize_beta=False, k_m=0.0, k_p=0.0):
    """Callbacks of a problem with spatially varying mu: alpha = u+ - u-,
    beta = -(mu+ grad u+ - mu- grad u-) . grad phi (the reference's sign; the
    level-set gradient as is, or normalised), g = u+ (the box lies in
    Omega+)."""
    phi = perturb(phi_raw)

    def alpha(r):
        return up(r) - um(r)

    def beta(r):
        g = grad_phi(r)
        if normalize_beta:
            nn = xnp.sqrt(g[0] ** 2 + g[1] ** 2 + g[2] ** 2)
            g = (g[0] / nn, g[1] / nn, g[2] / nn)
        a, b = grad_up(r), grad_um(r)
        mp, mm = mu_p(r), mu_m(r)
        v = sum((mp * a[i] - mm * b[i]) * g[i] for i in range(3))
        return xnp.nan_to_num(-v)

    def exact(r):
        return xnp.where(phi(r) >= 0, up(r), um(r))

    return dict(name=name, initial_value_fn=_const(0.0), dirichlet_bc_fn=up, lvl_set_fn=phi, mu_m_fn=mu_m,
                mu_p_fn=mu_p, k_m_fn=_const(k_m), k_p_fn=_const(k_p), f_m_fn=f_m, f_p_fn=f_p, alpha_fn=alpha,
                beta_fn=beta, exact_fn=exact, exact_m_fn=um, exact_p_fn=up, grad_mu_m_fn=grad_mu_m,
                grad_mu_p_fn=grad_mu_p, grad_exact_m_fn=grad_um, grad_exact_p_fn=grad_up, phi_raw_fn=phi_raw,
                grad_phi_fn=grad_phi)


def ref_sphere():
    """The reference's variable-coefficient sphere (tests/confs/experiment_configs.py:21-186):
    phi = |x| - 0.5, u- = exp(z), u+ = sin y cos x, mu- = y^2 log(x + 2) + 4,
    mu+ = exp(-z), k = 0; f = -div(mu grad u) on each side (derived here:
    f- = -exp(z) mu-, f+ = 2 exp(-z) cos x sin y), alpha = u+ - u-, beta with
    the unnormalised grad phi (a unit vector for this phi), g = u+.
    mu- > mu+ wherever |x| <= 1 (the mu- > mu+ ghost-value branch)."""

    def phi(r):
        return xnp.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2) - 0.5

    def gphi(r):
        rr = xnp.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2)
        return (r[0] / rr, r[1] / rr, r[2] / rr)

    def um(r):
        return xnp.exp(r[2])

    def up(r):
        return xnp.sin(r[1]) * xnp.cos(r[0])

    def gum(r):
        z = xnp.zeros_like(r[0])
        return (z, z, xnp.exp(r[2]))

    def gup(r):
        return (-xnp.sin(r[1]) * xnp.sin(r[0]), xnp.cos(r[1]) * xnp.cos(r[0]), xnp.zeros_like(r[0]))

    def mu_m(r):
        return r[1] * r[1] * xnp.log(r[0] + 2.0) + 4.0

    def mu_p(r):
        return xnp.exp(-r[2])

    def gmu_m(r):
        return (r[1] * r[1] / (r[0] + 2.0), 2.0 * r[1] * xnp.log(r[0] + 2.0), xnp.zeros_like(r[0]))

    def gmu_p(r):
        z = xnp.zeros_like(r[0])
        return (z, z, -xnp.exp(-r[2]))

    def f_m(r):
        # -d/dz (mu- exp(z)); mu- does not depend on z
        return -mu_m(r) * xnp.exp(r[2])

    def f_p(r):
        # -(mu+ lap u+ + grad mu+ . grad u+) with lap u+ = -2 u+ and grad mu+ . grad u+ = 0
        return 2.0 * xnp.exp(-r[2]) * xnp.cos(r[0]) * xnp.sin(r[1])

    return _problem(phi, gphi, mu_m, mu_p, gmu_m, gmu_p, um, up, gum, gup, f_m, f_p, "ref_sphere")


# the reference star's level set: three angular modes (n, beta, theta) about the z axis
_STAR_R0 = 0.483
_STAR_MODES = ((3.0, 0.1, 0.5), (4.0, -0.1, 1.8), (7.0, 0.15, 0.0))


def ref_star(normalize_beta=False):
    """The reference's variable-coefficient star (tests/confs/experiment_configs.py:196-412):
    phi = |x| - r0 (1 + ((x^2 + y^2)/|x|^2)^2 sum_i b_i cos(n_i (atan2(y, x) - t_i)))
    (r0 = 0.483, modes (3, 0.1, 0.5), (4, -0.1, 1.8), (7, 0.15, 0)), not finite
    at the origin -> 0 before the perturbation; u- = sin 2x cos 2y exp(z),
    u+ = T5((y - x)/3) log(x + y + 3) cos z with T5 the Chebyshev polynomial
    16t^5 - 20t^3 + 5t, mu- = 10 (1 + 0.2 cos(2 pi (x + y)) sin(2 pi (x - y)) cos z),
    mu+ = 1, k = 0; f = -div(mu grad u) derived here; beta with the
    unnormalised grad phi as the reference writes it (normalize_beta=True
    divides by |grad phi|, which the manufactured solution needs to satisfy the
    flux jump), not finite values -> 0; g = u+.  mu- > mu+ everywhere."""
    r0 = _STAR_R0

    def _core(ps):
        c = 0.0
        dc = 0.0
        for n_i, b_i, t_i in _STAR_MODES:
            c = c + b_i * xnp.cos(n_i * (ps - t_i))
            dc = dc - b_i * n_i * xnp.sin(n_i * (ps - t_i))
        return c, dc

    def phi(r):
        x, y, z = r[0], r[1], r[2]
        rho2 = x ** 2 + y ** 2 + z ** 2
        c, _ = _core(xnp.arctan2(y, x))
        v = xnp.sqrt(rho2) - r0 * (1.0 + ((x ** 2 + y ** 2) / rho2) ** 2 * c)
        return xnp.nan_to_num(v)

    def gphi(r):
        x, y, z = r[0], r[1], r[2]
        q = x ** 2 + y ** 2
        rho2 = q + z ** 2
        rho = xnp.sqrt(rho2)
        s = q / rho2
        c, dc = _core(xnp.arctan2(y, x))
        # grad(s^2 c) = 2 s c grad s + s^2 c' grad psi with
        # grad s = (2x z^2, 2y z^2, -2 q z) / rho^4 and s^2 grad psi = q (-y, x, 0) / rho^4
        sa = 2.0 * s * c / rho2 ** 2
        pb = q * dc / rho2 ** 2
        gx = x / rho - r0 * (sa * 2.0 * x * z ** 2 - pb * y)
        gy = y / rho - r0 * (sa * 2.0 * y * z ** 2 + pb * x)
        gz = z / rho - r0 * (sa * (-2.0 * q * z))
        return (gx, gy, gz)

    def um(r):
        return xnp.sin(2.0 * r[0]) * xnp.cos(2.0 * r[1]) * xnp.exp(r[2])

    def gum(r):
        e = xnp.exp(r[2])
        return (2.0 * xnp.cos(2.0 * r[0]) * xnp.cos(2.0 * r[1]) * e,
                -2.0 * xnp.sin(2.0 * r[0]) * xnp.sin(2.0 * r[1]) * e,
                xnp.sin(2.0 * r[0]) * xnp.cos(2.0 * r[1]) * e)

    def _t(r):
        t = (r[1] - r[0]) / 3.0
        T = 16.0 * t ** 5 - 20.0 * t ** 3 + 5.0 * t
        dT = 80.0 * t ** 4 - 60.0 * t ** 2 + 5.0
        ddT = 320.0 * t ** 3 - 120.0 * t
        return T, dT, ddT

    def up(r):
        T, _, _ = _t(r)
        return T * xnp.log(r[0] + r[1] + 3.0) * xnp.cos(r[2])

    def gup(r):
        T, dT, _ = _t(r)
        w = r[0] + r[1] + 3.0
        L, C = xnp.log(w), xnp.cos(r[2])
        return (-dT / 3.0 * L * C + T * C / w, dT / 3.0 * L * C + T * C / w, -T * L * xnp.sin(r[2]))

    def _mu_parts(r):
        A = xnp.cos(2.0 * xnp.pi * (r[0] + r[1]))
        B = xnp.sin(2.0 * xnp.pi * (r[0] - r[1]))
        C = xnp.cos(r[2])
        return A, B, C

    def mu_m(r):
        A, B, C = _mu_parts(r)
        return 10.0 * (1.0 + 0.2 * A * B * C)

    def gmu_m(r):
        A, B, C = _mu_parts(r)
        dA = -2.0 * xnp.pi * xnp.sin(2.0 * xnp.pi * (r[0] + r[1]))  # d/dx = d/dy
        dB = 2.0 * xnp.pi * xnp.cos(2.0 * xnp.pi * (r[0] - r[1]))  # d/dx = -d/dy
        return (2.0 * (dA * B + A * dB) * C, 2.0 * (dA * B - A * dB) * C, -2.0 * A * B * xnp.sin(r[2]))

    def mu_p(r):
        return xnp.zeros_like(r[0]) + 1.0

    def gmu_p(r):
        z = xnp.zeros_like(r[0])
        return (z, z, z)

    def f_m(r):
        # lap u- = -7 u-
        g, gm = gum(r), gmu_m(r)
        return 7.0 * mu_m(r) * um(r) - (gm[0] * g[0] + gm[1] * g[1] + gm[2] * g[2])

    def f_p(r):
        # lap u+ = (2/9) T'' L C - 2 T C / w^2 - T L C (the mixed terms of d_xx and d_yy cancel)
        T, _, ddT = _t(r)
        w = r[0] + r[1] + 3.0
        L, C = xnp.log(w), xnp.cos(r[2])
        return -((2.0 / 9.0) * ddT * L * C - 2.0 * T * C / w ** 2 - T * L * C)

    return _problem(phi, gphi, mu_m, mu_p, gmu_m, gmu_p, um, up, gum, gup, f_m, f_p, "ref_star",
                    normalize_beta=normalize_beta)


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
        # the same cell list as compressed rows, for the device kernel (dips_atoms_eval)
        self.cell_start = np.zeros(self.nc ** 3 + 1, np.int32)
        np.cumsum(counts, out=self.cell_start[1:])
        self.cell_atoms = np.argsort(flat, kind="stable").astype(np.int32)
        self.atoms5 = np.ascontiguousarray(np.concatenate([self.c, self.r[:, None], self.q[:, None]], axis=1))
        self._dev = {}

    def _eval_device(self, r, what):
        """phi or charge of GPU positions r (3, m) through dips_atoms_eval."""
        import torch

        from . import _native as N

        key = ("kernel", str(r.device))
        if key not in self._dev:
            self._dev[key] = tuple(torch.as_tensor(a, device=r.device) for a in
                                   (self.atoms5, self.cell_start, self.cell_atoms))
        atoms, cs, ca = self._dev[key]
        pos = r.t().to(torch.float64).contiguous()
        m = int(pos.shape[0])
        out = torch.empty(m, dtype=torch.float64, device=r.device)
        N.check(N.lib().dips_atoms_eval(N.ptr(atoms), len(self.c), N.ptr(cs), N.ptr(ca), self.nc, self.lo, self.cell,
                                        self.cap, self.sigma, N.ptr(pos), m, N.ptr(out) if what == "phi" else None,
                                        N.ptr(out) if what != "phi" else None, N.current_stream()), "dips_atoms_eval")
        return out

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
        if xnp._is_torch(r) and r.is_cuda:
            return self._eval_device(r, what)
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


def lpbe_sphere(h, sigma=1.0, charge=1.0, centre=(0.0, 0.0, 0.0), half_width=2.5):
    """The reference's linearized Poisson-Boltzmann benchmark with a known answer
    (examples/benchmark_LPBE: one charged sphere, inputs/sphere/single.pqr,
    radius 1 and charge +1 at the centre of the box [-2.5, 2.5]^3, main.py:139-140).

    Dimensionless constants (units.py:1-16): eps- = alpha_1 = 2, eps+ = alpha_2 = 80,
    kappa = sqrt(8.4869 Is / alpha_2) with Is = 1e-5, k- = 0, k+ = alpha_2 kappa^2,
    omega = 24.  With the point charge treated apart (u- = potential - g in the
    sphere, g = omega / (4 pi alpha_1) q / |x - c|, coefficients.py:125-163):
    f- = f+ = 0 (coefficients.py:294-301), alpha = [u] = g (coefficients.py:241-246)
    and beta = alpha_1 (-grad g) . n with n the level-set normal from central
    differences of phi with the training spacing h (coefficients.py:174-194,
    :221-252; -(mu+ grad u+ - mu- grad u-) . n in the reference's sign), the
    Dirichlet value omega / (4 pi alpha_2) q exp(kappa (sigma - d)) / (d (1 + kappa
    sigma)) at distance d (coefficients.py:35-52).  Exact solution
    (coefficients.py:307-324): u- = omega / (4 pi sigma) (1 / (alpha_2 (1 + kappa
    sigma)) - 1 / alpha_1), constant, and u+ = the Dirichlet expression."""
    alpha_1, alpha_2, omega = 2.0, 80.0, 24.0
    kappa = (8.4869 * 1e-5 / alpha_2) ** 0.5
    cx, cy, cz = centre
    g_coeff = omega / (4.0 * np.pi * alpha_1)
    bc_coeff = omega / (4.0 * np.pi * alpha_2)

    def dist(r):
        return xnp.sqrt((r[0] - cx) ** 2 + (r[1] - cy) ** 2 + (r[2] - cz) ** 2)

    def phi_raw(r):
        return dist(r) - sigma

    phi = perturb(phi_raw)

    def g(r):
        return xnp.nan_to_num(g_coeff * charge / dist(r))

    def beta(r):
        # minus grad g = g_coeff q (x - c) / d^3, against the normal of phi from central differences of step h
        d = dist(r)
        mg = [xnp.nan_to_num(g_coeff * charge * (r[i] - c) / d ** 3) for i, c in enumerate(centre)]
        nrm = []
        for i in range(3):
            rp = [r[j] + (h if j == i else 0.0) for j in range(3)]
            rm = [r[j] - (h if j == i else 0.0) for j in range(3)]
            nrm.append((phi(rp) - phi(rm)) / (2.0 * h))
        nn = xnp.sqrt(nrm[0] ** 2 + nrm[1] ** 2 + nrm[2] ** 2)
        return alpha_1 * sum(mg[i] * nrm[i] / nn for i in range(3))

    def up(r):
        d = dist(r)
        return xnp.nan_to_num(bc_coeff * charge * xnp.exp(kappa * (sigma - d)) / (d * (1.0 + kappa * sigma)))

    u_in = omega / (4.0 * np.pi * sigma) * (1.0 / (alpha_2 * (1.0 + kappa * sigma)) - 1.0 / alpha_1)

    def um(r):
        return xnp.zeros_like(r[0]) + u_in

    def exact(r):
        return xnp.where(phi(r) >= 0, up(r), um(r))

    zero = _const(0.0)
    return dict(name="lpbe", initial_value_fn=zero, dirichlet_bc_fn=up, lvl_set_fn=phi, mu_m_fn=_const(alpha_1),
                mu_p_fn=_const(alpha_2), k_m_fn=_const(0.0), k_p_fn=_const(alpha_2 * kappa ** 2), f_m_fn=zero,
                f_p_fn=zero, alpha_fn=g, beta_fn=beta, exact_fn=exact, exact_m_fn=um, exact_p_fn=up,
                phi_raw_fn=phi_raw, kappa=kappa, half_width=half_width)


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
