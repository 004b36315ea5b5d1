"""Per-point cut-cell finite-volume residual as an affine map of network outputs.

Test infrastructure only (see oracle/__init__.py).

Restates jax_dips/solvers/poisson/discretization.py:
  * get_Xijk (27-cube offsets, x fastest) .......... :164-197
  * normal_point_fn (central differences of phi) ... :199-218
  * get_regression_coeffs_at_point (pinv, zeta, gamma) :238-296
  * compute_Ax_and_b_preconditioned_fn ............. :299-423
  * get_u_mp_by_regression_at_point_fn ............. :426-519

For every training point p the reference residual
    r(p) = nan_to_num(lhs/diag) - nan_to_num(rhs/diag)
is affine in the network outputs, because geometry, regression rows and the
user callbacks do not depend on the parameters (only U = NN(position) does).
``point_slots`` returns that map in the reference's own structure: 7 stencil
points s_e (centre, x-, x+, y-, y+, z-, z+) times 27 cube slots q (slot 13 is
NN(s_e) itself), i.e.  r = sum_{e,q} W[e,q] * NN(s_e + X_q) + c.
``merge_slots`` folds slots that land on the same lattice offset
o = e + x_q in [-2,2]^3 into the 125-entry form the device records use.

Pinned by tests/test_oracle_discretization.py: the uncut-cell stencil equals
the textbook 7-point FV stencil [(6 mu h + k h^3) u_c - mu h sum u_nb - f h^3] /
(6 mu h + k h^3); box points give u - g; all-weights regression gives
D = X^T/(18 h^2); mu- = mu+ removes the regression terms so the ghost value is
u_c -+ (alpha + delta beta/mu); linear u is reproduced exactly by the ghost
extrapolation with equal mu; pinv cutoff matches 10*max(M,N)*eps_f32.
"""
import numpy as np

from . import geometry as G

# stencil points: centre, x-, x+, y-, y+, z-, z+ (discretization.py:347-353)
STENCIL = np.array([[0, 0, 0], [-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]])
# 27-cube offsets, q = 9(k+1) + 3(j+1) + (i+1) (discretization.py:164-197)
CUBE = np.array([[i, j, k] for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1)])
# neighbour q of the cube used by normal_point_fn: (+x, -x, +y, -y, +z, -z)
NORMAL_Q = (14, 12, 16, 10, 22, 4)
PINV_RCOND_F32 = 10.0 * 3 * float(np.finfo(np.float32).eps)
# local lattice offsets o in [-2,2]^3, code = 25(k+2) + 5(j+2) + (i+2)
OFFS125 = np.array([[i, j, k] for k in range(-2, 3) for j in range(-2, 3) for i in range(-2, 3)])


def offset_code(o):
    o = np.asarray(o)
    return 25 * (o[..., 2] + 2) + 5 * (o[..., 1] + 2) + (o[..., 0] + 2)


def f32_shift(P, off, h):
    """fl32(P + off*h) per axis, the way the reference builds stencil points."""
    P = np.asarray(P, dtype=np.float32)
    h32 = np.asarray(h, dtype=np.float32)
    return (P + (np.asarray(off, dtype=np.float32) * h32)).astype(np.float32)


class Callbacks:
    """Vectorised callbacks: each maps positions (m,3) -> values (m,) in dtype dt.

    ``fns`` holds user callbacks in coordinate-first form f(r) with r of shape
    (3, m) (so r[0], r[1], r[2] are coordinate vectors)."""

    NAMES = ("phi", "mu_m", "mu_p", "k_m", "k_p", "f_m", "f_p", "alpha", "beta", "g")

    def __init__(self, fns, dt):
        self.fns = fns
        self.dt = dt

    def __call__(self, name, R):
        R = np.asarray(R, dtype=self.dt)
        shape = R.shape[:-1]
        flat = R.reshape(-1, 3)
        with np.errstate(all="ignore"):
            v = self.fns[name](flat.T)
        v = np.broadcast_to(np.asarray(v, dtype=self.dt), (flat.shape[0],))
        return v.reshape(shape)


def box_boundary(P, h, bounds):
    """is_box_boundary_point (discretization.py:319-333), float32 arithmetic.
    bounds = (xmin, xmax, ymin, ymax, zmin, zmax) of the level-set grid."""
    P = np.asarray(P, dtype=np.float32)
    h = np.asarray(h, dtype=np.float32)
    b = np.asarray(bounds, dtype=np.float32)
    on = np.zeros(P.shape[0], dtype=bool)
    for a in range(3):
        tol = np.float32(1e-6) * h[a]
        on |= np.abs(P[:, a] - b[2 * a]) < tol
        on |= np.abs(P[:, a] - b[2 * a + 1]) < tol
    return on


def regression_rows(phi_cube, h, dt):
    """Least-squares gradient operators D- and D+ (3x27 each) from phi on the
    27-cube (discretization.py:249-260).  W+ = 1[phi>0], W- = 1[phi<0]."""
    X = (CUBE * np.asarray(h, dtype=np.float32)[None, :]).astype(np.float32).astype(dt)
    out = []
    for mask in (phi_cube < 0, phi_cube > 0):
        Wx = X[None, :, :] * mask[:, :, None].astype(dt)  # (n,27,3) = W X
        M = np.einsum("qa,nqb->nab", X, Wx)  # X^T W X
        with np.errstate(all="ignore"):
            Pi = np.linalg.pinv(M, rcond=dt(PINV_RCOND_F32))
            D = np.nan_to_num(Pi @ np.swapaxes(Wx, 1, 2))
        out.append(D.astype(dt))
    return out[0], out[1]  # Dm, Dp


def u_mp_forms(s, h, cb, dt):
    """Affine forms of (u-, u+) at stencil points s (n,3) in terms of the 27
    cube slots U_q = NN(s + X_q) (slot 13 is NN(s)).

    Returns a_m, a_p (n,27), b_m, b_p (n,) and aux data
    (get_u_mp_by_regression_at_point_fn, discretization.py:426-519)."""
    n = s.shape[0]
    corners = G.cell_corners(s, h, np.float32)
    region = G.region_from_corners(cb("phi", corners))
    delta = cb("phi", s)
    a_m = np.zeros((n, 27), dtype=dt)
    a_p = np.zeros((n, 27), dtype=dt)
    b_m = np.zeros(n, dtype=dt)
    b_p = np.zeros(n, dtype=dt)
    a_m[:, 13] = np.where(region == -1, 1, 0)
    a_p[:, 13] = np.where(region == 1, 1, 0)
    cut = np.nonzero(region == 0)[0]
    aux = dict(region=region, delta=delta, cut=cut)
    if cut.size == 0:
        return a_m, a_p, b_m, b_p, aux
    sc = s[cut]
    dc = delta[cut]
    cube_pos = (sc[:, None, :] + (CUBE * np.asarray(h, np.float32)[None, :]).astype(np.float32)[None]).astype(np.float32)
    phi_cube = cb("phi", cube_pos)
    Dm, Dp = regression_rows(phi_cube, h, dt)
    hh = np.asarray(h, dtype=dt)
    with np.errstate(all="ignore"):
        gx = (phi_cube[:, 14] - phi_cube[:, 12]) / (dt(2) * hh[0])
        gy = (phi_cube[:, 16] - phi_cube[:, 10]) / (dt(2) * hh[1])
        gz = (phi_cube[:, 22] - phi_cube[:, 4]) / (dt(2) * hh[2])
        nrm = np.sqrt(gx * gx + gy * gy + gz * gz)
        nvec = np.stack([gx / nrm, gy / nrm, gz / nrm], axis=1)
    mu_m_s = cb("mu_m", sc)
    mu_p_s = cb("mu_p", sc)
    Cm = np.einsum("na,naq->nq", nvec, Dm)
    Cp = np.einsum("na,naq->nq", nvec, Dp)
    with np.errstate(all="ignore"):
        zp_q = (((mu_p_s - mu_m_s) / mu_m_s) * dc)[:, None] * Cp
        zm_q = (((mu_p_s - mu_m_s) / mu_p_s) * dc)[:, None] * Cm
        zp = -(zp_q.sum(axis=1) - zp_q[:, 13])
        zm = -(zm_q.sum(axis=1) - zm_q[:, 13])
        gp_q = zp_q / (dt(1) + zp)[:, None]
        gm_q = zm_q / (dt(1) - zm)[:, None]
        gp = -(gp_q.sum(axis=1) - gp_q[:, 13])
        gm = -(gm_q.sum(axis=1) - gm_q[:, 13])
        r_proj = sc.astype(dt) - dc[:, None] * nvec
        al = cb("alpha", r_proj)
        be = cb("beta", r_proj)
        jump_p = al + dc * (be / cb("mu_p", r_proj))  # alpha + delta*beta/mu+
        jump_m = al + dc * (be / cb("mu_m", r_proj))  # alpha + delta*beta/mu-

    def extrap(row_q, row_s, const):
        # -row.U + (1 - row_s + row_q[13]) u_c + const, with u_c = U_13
        a = -row_q.copy()
        a[:, 13] = a[:, 13] + (dt(1) - row_s + row_q[:, 13])
        return a, const

    one_c = np.zeros((cut.size, 27), dtype=dt)
    one_c[:, 13] = 1
    mbig = mu_m_s > mu_p_s
    pos = dc > 0
    # mu- > mu+ (:465-486)
    am1, bm1 = extrap(gm_q, gm, -(dt(1) - gm) * jump_p)
    ap1, bp1 = extrap(zm_q, zm, jump_p)
    # otherwise (:488-509)
    am2, bm2 = extrap(zp_q, zp, -jump_m)
    ap2, bp2 = extrap(gp_q, gp, (dt(1) - gp) * jump_m)
    zero = np.zeros(cut.size, dtype=dt)
    A_m = np.where(mbig[:, None], np.where(pos[:, None], am1, one_c), np.where(pos[:, None], am2, one_c))
    B_m = np.where(mbig, np.where(pos, bm1, zero), np.where(pos, bm2, zero))
    A_p = np.where(mbig[:, None], np.where(pos[:, None], one_c, ap1), np.where(pos[:, None], one_c, ap2))
    B_p = np.where(mbig, np.where(pos, zero, bp1), np.where(pos, zero, bp2))
    a_m[cut] = A_m
    a_p[cut] = A_p
    b_m[cut] = B_m
    b_p[cut] = B_p
    aux.update(nvec=nvec, r_proj=r_proj, Dm=Dm, Dp=Dp)
    return a_m, a_p, b_m, b_p, aux


def point_slots(P, h, bounds, cb, dt=np.float64):
    """Residual of every point as r = sum_{e,q} W[e,q] NN(s_e + X_q) + c.

    P: (n,3) float32 training points, h: (3,) float32 stencil spacing,
    bounds: level-set grid bounds.  Returns dict W (n,7,27), c (n,), boundary
    (n,), dead (n,) [lhs/diag not finite -> nan_to_num gives 0], plus the
    float32 slot positions pos (n,7,27,3).
    (compute_Ax_and_b_preconditioned_fn, discretization.py:299-423; the
    preconditioner is disabled, P = 1, trainer.py:246-253.)"""
    P = np.asarray(P, dtype=np.float32)
    n = P.shape[0]
    hh = np.asarray(h, dtype=dt)
    vol = hh[0] * hh[1] * hh[2]
    bnd = box_boundary(P, h, bounds)

    geo = G.cell_geometry(P, h, cb("phi", G.cell_corners(P, h, np.float32)), dt)
    fcs = G.face_centres(P, h, np.float32)
    mu_m_f = cb("mu_m", fcs)
    mu_p_f = cb("mu_p", fcs)
    hax = hh[G.FACE_AXIS][None, :]
    c_m = geo["area_m"] * mu_m_f / hax
    c_p = geo["area_p"] * mu_p_f / hax
    V_m, V_p = geo["vol_m"], geo["vol_p"]
    k_m, k_p = cb("k_m", P), cb("k_p", P)
    sum_cm = c_m[:, 0] + c_m[:, 1] + c_m[:, 2] + c_m[:, 3] + c_m[:, 4] + c_m[:, 5]
    sum_cp = c_p[:, 0] + c_p[:, 1] + c_p[:, 2] + c_p[:, 3] + c_p[:, 4] + c_p[:, 5]
    diag = k_p * V_p + k_m * V_m + sum_cm + sum_cp
    with np.errstate(all="ignore"):
        rhs = cb("f_m", P) * V_m + cb("f_p", P) * V_p
        rhs = rhs + G.interface_integral(geo, lambda R: cb("beta", R), dt)

    W = np.zeros((n, 7, 27), dtype=dt)
    lhs_c = np.zeros(n, dtype=dt)
    pos = np.zeros((n, 7, 27, 3), dtype=np.float32)
    regions = np.zeros((n, 7), dtype=np.int64)
    for e in range(7):
        s = f32_shift(P, STENCIL[e], h)
        pos[:, e] = (s[:, None, :] + (CUBE * np.asarray(h, np.float32)[None, :]).astype(np.float32)[None]).astype(np.float32)
        a_m, a_p, b_m, b_p, aux = u_mp_forms(s, h, cb, dt)
        regions[:, e] = aux["region"]
        if e == 0:
            cm, cp = k_m * V_m + sum_cm, k_p * V_p + sum_cp
        else:
            cm, cp = -c_m[:, e - 1], -c_p[:, e - 1]
        with np.errstate(all="ignore"):
            W[:, e] = cm[:, None] * a_m + cp[:, None] * a_p
            lhs_c = lhs_c + cm * b_m + cp * b_p
    with np.errstate(all="ignore"):
        Wd = W / diag[:, None, None]
        lc = lhs_c / diag
        rd = rhs / diag
    dead = ~(np.isfinite(Wd).all(axis=(1, 2)) & np.isfinite(lc))
    Wd[dead] = 0
    lc = np.where(dead, dt(0), lc)
    rd = np.nan_to_num(rd, nan=0.0, posinf=np.finfo(np.float32).max, neginf=-np.finfo(np.float32).max).astype(dt)
    c = lc - rd
    # box points: (u h^3 / h^3) - (g h^3 / h^3)
    g = cb("g", P)
    Wd[bnd] = 0
    Wd[bnd, 0, 13] = 1
    c = np.where(bnd, -np.nan_to_num(g * vol / vol), c)
    dead = dead & ~bnd
    return dict(W=Wd, c=c, boundary=bnd, dead=dead, pos=pos, regions=regions, geo=geo,
                c_m=c_m, c_p=c_p, diag=diag, rhs=rhs)


def merge_slots(slots, P, h, phi_of):
    """Fold (stencil, cube) slots onto (lattice offset, phase) keys.

    A slot's network phase is phi(s_e + X_q) >= 0 at its own float32 position
    (MLP.py:98); slots landing on the same offset o = e + x_q in [-2,2]^3 with
    the same phase are summed.  Returns W250 (n,250) indexed 2*code + phase and
    the canonical float32 positions fl32(p + o*h) (n,125,3) at which the folded
    entry is evaluated."""
    W = slots["W"]
    n = W.shape[0]
    slot_phase = phi_of(slots["pos"]) >= 0  # (n,7,27)
    W250 = np.zeros((n, 250), dtype=W.dtype)
    for e in range(7):
        for q in range(27):
            code = offset_code(STENCIL[e] + CUBE[q])
            key = 2 * code + slot_phase[:, e, q].astype(np.int64)
            np.add.at(W250, (np.arange(n), key), W[:, e, q])
    pos = f32_shift(np.asarray(P, np.float32)[:, None, :], OFFS125[None], h)
    return W250, pos
