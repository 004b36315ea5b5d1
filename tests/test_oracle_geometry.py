"""Pins the oracle's Min-Gibou geometry (oracle/geometry.py).

* Linear phi: the 5-tet decomposition is exact, so V-, the per-face apertures
  A- and the interface area must equal an independent exact computation: the
  convex polytope {phi < 0} cut from the cell (scipy ConvexHull volume), the
  convex polygon {phi < 0} on each face and the plane-cube polygon.
* Invariants: V- + V+ = h^3, A- + A+ = face area, phi -> -phi swaps the sides.
* The reference's own known-answer test: sphere r = 0.5 on [-2,2]^3 at 128^3,
  sum of interface areas = pi and sum of V- = pi/6 within 0.02
  (tests/test_geometric_integrations.py:172-182, 215-221 of the reference).
"""
import numpy as np
import pytest
from scipy.spatial import ConvexHull

from oracle import geometry as G

CUBE_EDGES = [(a, b) for a in range(8) for b in range(a + 1, 8)
              if np.abs(G.CORNER_SIGNS[a] - G.CORNER_SIGNS[b]).sum() == 2]


def exact_plane_cut(c, h, nrm, d):
    """phi(x) = nrm.x + d on the cell centred at c with sides h."""
    corners = c + G.CORNER_SIGNS * (h / 2)
    phi = corners @ nrm + d
    pts = [corners[i] for i in range(8) if phi[i] < 0]
    cross = []
    for a, b in CUBE_EDGES:
        if (phi[a] < 0) != (phi[b] < 0):
            t = phi[a] / (phi[a] - phi[b])
            cross.append(corners[a] + t * (corners[b] - corners[a]))
    vol = 0.0
    allp = np.array(pts + cross)
    if len(pts) == 8:
        vol = np.prod(h)
    elif len(allp) >= 4:
        try:
            vol = ConvexHull(allp).volume
        except Exception:
            vol = 0.0
    area = 0.0
    if len(cross) >= 3:
        u = np.cross(nrm, [1.0, 0.3, 0.1])
        u /= np.linalg.norm(u)
        v = np.cross(nrm / np.linalg.norm(nrm), u)
        P2 = np.array([[p @ u, p @ v] for p in cross])
        area = ConvexHull(P2).volume
    faces = []
    for f in range(6):
        ax, sg = G.FACE_AXIS[f], G.FACE_SIGN[f]
        others = [a for a in range(3) if a != ax]
        fc = c[ax] + sg * h[ax] / 2
        sq = []
        for s0 in (-1, 1):
            for s1 in (-1, 1):
                p = c.copy()
                p[ax] = fc
                p[others[0]] += s0 * h[others[0]] / 2
                p[others[1]] += s1 * h[others[1]] / 2
                sq.append(p)
        sq = np.array(sq)
        ph = sq @ nrm + d
        pts2 = [sq[i] for i in range(4) if ph[i] < 0]
        for a, b in ((0, 1), (0, 2), (1, 3), (2, 3)):
            if (ph[a] < 0) != (ph[b] < 0):
                t = ph[a] / (ph[a] - ph[b])
                pts2.append(sq[a] + t * (sq[b] - sq[a]))
        if sum(ph < 0) == 4:
            faces.append(h[others[0]] * h[others[1]])
        elif len(pts2) >= 3:
            try:
                faces.append(ConvexHull(np.array(pts2)[:, others]).volume)
            except Exception:
                faces.append(0.0)
        else:
            faces.append(0.0)
    return vol, np.array(faces), area


def random_planes(n, seed):
    rng = np.random.default_rng(seed)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    d = rng.uniform(-0.8, 0.8, size=n) * 0.5
    return nrm, d


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_plane_cuts_exact(seed):
    # dyadic centre and spacing: the float32 corner positions the oracle forms
    # (as the reference does) are then exact
    h = np.array([0.25, 0.25, 0.25])
    c = np.array([0.125, -0.375, 0.25])
    nrm, d = random_planes(60, seed)
    d = d * 0.35 - (nrm @ c)  # planes through the cell
    P = np.repeat(c[None], len(d), axis=0)
    corners = G.cell_corners(P, h, np.float64)
    phi_c = np.einsum("nka,na->nk", corners, nrm) + d[:, None]
    geo = G.cell_geometry(P, h, phi_c)
    for i in range(len(d)):
        vol, faces, area = exact_plane_cut(c, h, nrm[i], d[i])
        assert abs(geo["vol_m"][i] - vol) <= 1e-10, i
        assert np.allclose(geo["area_m"][i], faces, atol=1e-10), (i, geo["area_m"][i], faces)
        if geo["region"][i] == 0:
            assert abs(geo["tri_area"][i].sum() - area) <= 1e-10, i


def test_invariants_and_symmetry():
    h = np.array([0.125, 0.1875, 0.15625])  # dyadic: float32 corners are exact
    rng = np.random.default_rng(7)
    P = np.round(rng.uniform(-0.5, 0.5, size=(400, 3)) * 1024) / 1024
    f = lambda R: np.sqrt(((R - 0.05) ** 2).sum(-1)) - 0.37
    corners = G.cell_corners(P, h, np.float64)
    phi = f(corners)
    g1 = G.cell_geometry(P, h, phi)
    g2 = G.cell_geometry(P, h, -phi)
    vol = h.prod()
    face = np.array([h[1] * h[2]] * 2 + [h[0] * h[2]] * 2 + [h[0] * h[1]] * 2)
    assert np.allclose(g1["vol_m"] + g1["vol_p"], vol, rtol=1e-12, atol=1e-15)
    assert np.allclose(g1["area_m"] + g1["area_p"], face[None], rtol=1e-12, atol=1e-15)
    assert np.allclose(g1["vol_m"], g2["vol_p"], atol=1e-14)
    assert np.allclose(g1["area_m"], g2["area_p"], atol=1e-14)
    cut = g1["region"] == 0
    assert cut.sum() > 20
    assert np.allclose(g1["tri_area"][cut].sum(1), g2["tri_area"][cut].sum(1), atol=1e-14)
    assert np.array_equal(g1["region"], -g2["region"])


def test_brute_force_volume_curved():
    """Curved interface: V- converges to the sub-sampled volume as the cut
    becomes locally planar (second-order agreement on a small cell)."""
    h = np.array([0.015625, 0.015625, 0.015625])
    c = np.array([0.3125, 0.171875, 0.21875])
    f = lambda R: np.sqrt((R ** 2).sum(-1)) - 0.41
    geo = G.cell_geometry(c[None], h, f(G.cell_corners(c[None], h, np.float64)))
    m = 60
    g = (np.arange(m) + 0.5) / m - 0.5
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pts = c + np.stack([X, Y, Z], -1).reshape(-1, 3) * h
    brute = (f(pts) < 0).mean() * h.prod()
    assert geo["region"][0] == 0
    assert abs(geo["vol_m"][0] - brute) <= 2e-3 * h.prod()


def test_reference_sphere_known_answer():
    """Reference test: sum of interface areas ~ pi and sum of V- ~ pi/6 (0.02),
    sphere r = 0.5 on [-2,2]^3 at 128^3."""
    n = 128
    x = np.linspace(-2, 2, n, dtype=np.float32)
    h = np.array([x[1] - x[0]] * 3, np.float32)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    f = lambda R: np.sqrt((np.asarray(R, np.float64) ** 2).sum(-1)) - 0.5
    near = np.abs(f(P)) < 2 * h[0]
    region_far = np.where(f(P) < 0, -1, 1)
    Pn = P[near]
    geo = G.cell_geometry(Pn, h, f(G.cell_corners(Pn, h, np.float32)))
    area = geo["tri_area"][geo["region"] == 0].sum()
    vol = geo["vol_m"].sum() + np.float64(h.prod()) * ((region_far == -1) & ~near).sum()
    assert abs(area - np.pi) <= 0.02, area
    assert abs(vol - np.pi / 6) <= 0.02, vol
