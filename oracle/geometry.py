"""Min & Gibou (2007) cut-cell geometry around a point — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates jax_dips/geometry/geometric_integrations_per_point.py:
  * corner order and 5-tet split ............ :276-316
  * crossing test / region .................. :203-263
  * Gamma pieces per tet (eta = 1, 2, 3) ..... :52-107
  * Omega^- pieces per tet ................... :110-198
  * vol_fn / area_fn (Gram determinants) ..... :370-380, :496-507
  * volumes, face apertures, mu*A/h ......... :514-903
  * uncut cells ............................. :905-996
  * interface integral of beta .............. :383-420

Everything is vectorised over ``n`` cells and computed in the dtype ``dt``
(float64 for the oracle proper, float32 to mimic the reference's arithmetic).

Pinned by tests/test_oracle_geometry.py: plane cuts (exact volumes, face
apertures and interface areas for linear phi), V- + V+ = h^3, A- + A+ = face
area, phi -> -phi symmetry, sub-sampled brute-force volume on random planes,
and the reference's sphere known-answer test (sum of interface areas = pi,
sum of V- = pi/6 within 0.02 at 128^3 on [-2,2]^3;
tests/test_geometric_integrations.py:172-182, 215-221).
"""
import numpy as np

# P000, P100, P101, P001, P010, P110, P011, P111 (per-point file :276-288)
CORNER_SIGNS = np.array(
    [[-1, -1, -1], [1, -1, -1], [1, -1, 1], [-1, -1, 1],
     [-1, 1, -1], [1, 1, -1], [-1, 1, 1], [1, 1, 1]], dtype=np.int64)
# S1..S5 as corner indices (:312-316)
TETS = np.array([[0, 1, 4, 3], [5, 1, 4, 7], [2, 1, 7, 3], [6, 7, 4, 3], [7, 1, 4, 3]])
# Omega^- pieces feeding each face x-, x+, y-, y+, z-, z+ (:841-854)
FACE_TETS = np.array([[0, 3], [1, 2], [0, 2], [1, 3], [0, 1], [2, 3]])
FACE_AXIS = np.array([0, 0, 1, 1, 2, 2])
FACE_SIGN = np.array([-1, 1, -1, 1, -1, 1])


def cell_corners(P, h, dt):
    """Corner positions p + 0.5*[+-h], formed in float32 as the reference forms
    them (per-point file :276-290), returned in ``dt``."""
    P = np.asarray(P, dtype=np.float32)
    half = (np.asarray(h, dtype=np.float32) * np.float32(0.5))
    off = CORNER_SIGNS.astype(np.float32) * half[None, :]
    return (P[:, None, :] + off[None, :, :]).astype(dt)


def region_from_corners(phi_c):
    """0 if crossed, else sign(phi(P000)) (:261-263).  phi_c: (n, 8)."""
    eta = (phi_c < 0).sum(axis=1)
    crossed = (eta > 0) & (eta < 8)
    return np.where(crossed, 0, np.sign(phi_c[:, 0])).astype(np.int64)


def _gram_measure(A, dt):
    """sqrt(|det((A[1:]-A[0]) (A[1:]-A[0])^T)|) for stacks of simplices (:370-380)."""
    M = A[..., 1:, :] - A[..., :1, :]
    G = M @ np.swapaxes(M, -1, -2)
    d = np.linalg.det(G.astype(np.float64)).astype(dt) if dt == np.float64 else _det_f32(G)
    d = np.nan_to_num(d)
    return np.sqrt(np.abs(d))


def _det_f32(G):
    # determinant evaluated in float32 (cofactor expansion); the reference runs
    # jnp.linalg.det (LU) in float32 -- both are float32-rounded, neither exact.
    if G.shape[-1] == 3:
        a, b, c = G[..., 0, 0], G[..., 0, 1], G[..., 0, 2]
        d, e, f = G[..., 1, 0], G[..., 1, 1], G[..., 1, 2]
        g, h, i = G[..., 2, 0], G[..., 2, 1], G[..., 2, 2]
        return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    return G[..., 0, 0] * G[..., 1, 1] - G[..., 0, 1] * G[..., 1, 0]


def tet_volume(A, dt):
    return _gram_measure(A, dt) * dt(1.0 / 6.0)


def tri_area(A, dt):
    return _gram_measure(A, dt) * dt(0.5)


def _edge(Ss, ps, a, b):
    """Q(a,b) = (phi_a*S_b - phi_b*S_a)/(phi_a - phi_b) on sorted vertices."""
    pa = ps[:, a][:, None]
    pb = ps[:, b][:, None]
    with np.errstate(divide="ignore", invalid="ignore"):
        return (pa * Ss[:, b] - pb * Ss[:, a]) / (pa - pb)


def tet_pieces(S, phi, dt):
    """Gamma triangles (n,2,3,3) and Omega^- sub-tets (n,3,4,3) of one tet.

    S: (n,4,3) vertices, phi: (n,4).  Follows :52-198: vertices are sorted by
    phi with a stable sort; eta = #(phi < 0) selects the case; eta in {0, 4}
    give zero Gamma pieces; eta = 4 keeps the tet in its original vertex order.
    """
    n = S.shape[0]
    eta = (phi < 0).sum(axis=1)
    order = np.argsort(phi, axis=1, kind="stable")
    Ss = np.take_along_axis(S, order[:, :, None], axis=1)
    ps = np.take_along_axis(phi, order, axis=1)
    order_n = np.argsort(-phi, axis=1, kind="stable")
    Sn = np.take_along_axis(S, order_n[:, :, None], axis=1)
    pn = np.take_along_axis(-phi, order_n, axis=1)

    Q01, Q02, Q03 = _edge(Ss, ps, 0, 1), _edge(Ss, ps, 0, 2), _edge(Ss, ps, 0, 3)
    Q12, Q13, Q23 = _edge(Ss, ps, 1, 2), _edge(Ss, ps, 1, 3), _edge(Ss, ps, 2, 3)
    N01, N02, N03 = _edge(Sn, pn, 0, 1), _edge(Sn, pn, 0, 2), _edge(Sn, pn, 0, 3)
    S0, S1, S2 = Ss[:, 0], Ss[:, 1], Ss[:, 2]
    Z = np.zeros((n, 3), dtype=dt)

    def stack_tets(rows):
        return np.stack([np.stack(r, axis=1) for r in rows], axis=1)

    g1 = stack_tets([[Q01, Q02, Q03], [Z, Z, Z]])
    g2 = stack_tets([[Q02, Q03, Q13], [Q02, Q12, Q13]])
    g3 = stack_tets([[N01, N02, N03], [Z, Z, Z]])
    gz = np.zeros((n, 2, 3, 3), dtype=dt)
    e = eta[:, None, None, None]
    gamma = np.where(e == 1, g1, np.where(e == 2, g2, np.where(e == 3, g3, gz)))

    o1 = stack_tets([[S0, Q01, Q02, Q03], [Z, Z, Z, Z], [Z, Z, Z, Z]])
    o2 = stack_tets([[S0, S1, Q02, Q13], [Q12, S1, Q02, Q13], [S0, Q03, Q02, Q13]])
    o3 = stack_tets([[S0, S1, S2, Q13], [S0, Q03, S2, Q13], [Q23, Q03, S2, Q13]])
    o4 = np.zeros((n, 3, 4, 3), dtype=dt)
    o4[:, 0] = S
    oz = np.zeros((n, 3, 4, 3), dtype=dt)
    omega = np.where(e == 1, o1, np.where(e == 2, o2, np.where(e == 3, o3, np.where(e == 4, o4, oz))))
    return gamma, omega


def _isclose(a, b, atol):
    # jnp.isclose default rtol = 1e-5 (per-point file :812-833)
    return np.abs(a - b) <= (atol + 1e-5 * np.abs(b))


def cell_geometry(P, h, phi_c, dt=np.float64):
    """Full geometry of the cells of size h centred at P.

    Returns dict with region (n,), vol_m, vol_p (n,), area_m, area_p (n,6) in
    face order x-, x+, y-, y+, z-, z+, tri (n,10,3,3) and tri_area (n,10) in the
    reference's summation order S1 tri0, S1 tri1, ..., S5 tri1.
    """
    P = np.asarray(P, dtype=dt)
    h = np.asarray(h, dtype=dt)
    phi_c = np.asarray(phi_c, dtype=dt)
    n = P.shape[0]
    C = cell_corners(P, h, dt)
    region = region_from_corners(phi_c)

    gam = []
    omg = []
    for t in range(5):
        g, o = tet_pieces(C[:, TETS[t]], phi_c[:, TETS[t]], dt)
        gam.append(g)
        omg.append(o)
    omega = np.stack(omg, axis=1)  # (n,5,3,4,3)
    vols = tet_volume(omega, dt)  # (n,5,3)
    vol_m = np.zeros(n, dtype=dt)
    for t in range(5):
        for k in range(3):
            vol_m = vol_m + vols[:, t, k]
    vol = h[0] * h[1] * h[2]
    vol_p_cut = np.where(vol - vol_m < 0, dt(0), vol - vol_m)

    face_full = np.array([h[1] * h[2], h[1] * h[2], h[0] * h[2], h[0] * h[2], h[0] * h[1], h[0] * h[1]], dtype=dt)
    area_m_cut = np.zeros((n, 6), dtype=dt)
    for f in range(6):
        ax = FACE_AXIS[f]
        fc = (np.asarray(P[:, ax], np.float32) + np.float32(FACE_SIGN[f]) * (np.float32(0.5) * np.float32(h[ax]))).astype(dt)
        tot = np.zeros(n, dtype=dt)
        for t in FACE_TETS[f]:
            for k in range(3):
                verts = omega[:, t, k]  # (n,4,3)
                on = _isclose(verts[:, :, ax], fc[:, None], dt(1e-10) * h[ax])
                cnt = on.sum(axis=1)
                # the three on-face vertices, in index order (area is order independent)
                idx = np.argsort(~on, axis=1, kind="stable")[:, :3]
                tri = np.take_along_axis(verts, idx[:, :, None], axis=1)
                a = tri_area(tri, dt)
                tot = tot + np.where(cnt == 3, a, dt(0))
        area_m_cut[:, f] = tot
    area_p_cut = np.where(face_full[None, :] - area_m_cut < 0, dt(0), face_full[None, :] - area_m_cut)

    tri = np.concatenate(gam, axis=1)  # (n,10,3,3)
    tri_a = tri_area(tri, dt)

    cut = region == 0
    mmask = (region == -1).astype(dt)
    pmask = (region == 1).astype(dt)
    vol_m_out = np.where(cut, vol_m, vol * mmask)
    vol_p_out = np.where(cut, vol_p_cut, vol * pmask)
    area_m = np.where(cut[:, None], area_m_cut, face_full[None, :] * mmask[:, None])
    area_p = np.where(cut[:, None], area_p_cut, face_full[None, :] * pmask[:, None])
    return dict(region=region, vol_m=vol_m_out, vol_p=vol_p_out, area_m=area_m, area_p=area_p,
                tri=tri, tri_area=tri_a, corners=C)


def face_centres(P, h, dt):
    """p + 0.5*[+-h] along each axis, order x-, x+, y-, y+, z-, z+ (:533-545)."""
    P = np.asarray(P, dtype=dt)
    h = np.asarray(h, dtype=dt)
    out = np.repeat(P[:, None, :], 6, axis=1).copy()
    for f in range(6):
        ax = FACE_AXIS[f]
        out[:, f, ax] = P[:, ax] + FACE_SIGN[f] * (dt(0.5) * h[ax])
    return out


def interface_integral(geo, beta_at, dt):
    """Sum over the 10 triangle slots of area * mean(beta at 3 vertices), for
    crossed cells only (:383-420).  Zero (padding) slots sit at the origin, so a
    beta that is not finite there makes the integral NaN, as in the reference.
    ``beta_at`` maps (m,3) positions to (m,) values."""
    tri = geo["tri"]
    n = tri.shape[0]
    vals = beta_at(tri.reshape(-1, 3)).reshape(n, 10, 3).astype(dt)
    mean = (vals[:, :, 0] + vals[:, :, 1] + vals[:, :, 2]) / dt(3.0)
    acc = np.zeros(n, dtype=dt)
    for s in range(10):
        acc = acc + geo["tri_area"][:, s] * mean[:, s]
    return np.where(geo["region"] == 0, acc, dt(0))
