"""Build the device records of one zoom level with the FP64 geometry kernels.

Per zoom level (stencil spacing h) the reference recomputes, at every step and
for every training point, the cut-cell geometry of the 7 stencil cells, the
regression rows at crossed stencil points and the callbacks
(discretization.py:299-519).  None of it depends on the network parameters,
so it is evaluated once here:

  1. callback samples at the positions the reference uses, built in float32
     exactly as it builds them: s_e = p + e*h, corners s_e + (+-h/2),
     cube points s_e + X_q, face centres p + (+-h/2) e_a
     (geometric_integrations_per_point.py:276-290, :533-545; discretization.py:
     164-197, :347-353, :441);
  2. dips_geom_regions: crossing test of every stencil cell;
  3. dips_geom_regression at crossed stencil points -> affine ghost-value
     forms and the interface projection r_Gamma = s - delta n;
  4. dips_geom_cells for crossed centre cells -> V-, face apertures, Gamma
     triangles; then alpha/beta/mu at r_Gamma and beta at triangle vertices;
  5. dips_geom_assemble (count, then write) -> CSR rows ordered by entry count
     inside each batch.

Callbacks run on torch tensors on the GPU when they accept them (the
``xnp`` namespace) and on numpy arrays on the host otherwise; positions are
float32, samples float64.  Index bookkeeping (lists of crossed cells, the row
order and CSR offsets) uses torch device primitives.
"""
import ctypes

import numpy as np

from . import _native as N
from .records import BULK_MAX_ENTRIES, DeviceRecords

CORNER_SIGNS = ((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1), (-1, 1, -1), (1, 1, -1), (-1, 1, 1), (1, 1, 1))
STENCIL = ((0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1))
CUBE = tuple((i, j, k) for k in (-1, 0, 1) for j in (-1, 0, 1) for i in (-1, 0, 1))
FACES = ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1))
PINV_RCOND = 10.0 * 3 * float(np.finfo(np.float32).eps)
CHUNK = 1 << 24


def sample(fn, pos):
    """Evaluate a coordinate-first callback at positions pos (..., 3) -> float64 (...)."""
    import torch

    shape = pos.shape[:-1]
    flat = pos.reshape(-1, 3)
    m = flat.shape[0]
    out = torch.empty(m, dtype=torch.float64, device=pos.device)
    for a in range(0, m, CHUNK):
        b = min(m, a + CHUNK)
        r = flat[a:b].to(torch.float64).t()
        try:
            v = fn(r)
            if not torch.is_tensor(v):
                v = torch.as_tensor(np.asarray(v, dtype=np.float64), device=pos.device)
        except (TypeError, AttributeError, RuntimeError):
            v = torch.as_tensor(np.asarray(fn(r.cpu().numpy()), dtype=np.float64), device=pos.device)
        out[a:b] = v.to(torch.float64).expand(b - a)
    return out.reshape(shape)


def _offsets(vals, device):
    import torch

    return torch.tensor(vals, dtype=torch.float32, device=device)


def stencil_corner_phi(grid, S, ht, half, phi):
    """phi at the 7 stencil points s_e = fl32(p + e h) of every training point,
    (N,7), and at the 8 corners fl32(s_e + c h/2) of each stencil cell,
    (N,7,8), float64.

    For the grid nodes every coordinate of these positions depends on its own
    axis only (node index, stencil sign, corner sign), so the positions are
    formed from per-axis tables of the distinct float32 values, and phi is
    evaluated once per distinct position: neighbouring points share cells and
    corners (at the training spacing a node's 63 samples have about 5
    distinct positions, at h/8 about 45).  The values are the same floats the
    direct construction gives (same operations, same rounding).  Extra points
    (batch padding) are sampled directly."""
    import torch

    device = S.device
    nx, ny, nz = grid.shape
    ng = grid.ngrid
    if ng == 0:  # scattered points only: every sample taken directly
        phis = sample(phi, S).contiguous()
        phic = sample(phi, S[:, :, None, :] + _offsets(CORNER_SIGNS, device)[None, None] * half).contiguous()
        return phis, phic
    sgn = torch.tensor([-1.0, 0.0, 1.0], dtype=torch.float32, device=device)
    tabs = []
    for a, axv in enumerate((grid.xs, grid.ys, grid.zs)):
        V = axv[:, None] + sgn[None, :] * ht[a]  # stencil coordinate: node value + e h
        W = V[:, :, None] + sgn[None, None, :] * half[a]  # corner coordinate; sign 0 gives V itself
        U, inv = torch.unique(W.reshape(-1), sorted=True, return_inverse=True)
        tabs.append((U, inv.reshape(len(axv), 3, 3)))
    (Ux, Ix), (Uy, Iy), (Uz, Iz) = tabs
    My, Mz = Uy.numel(), Uz.numel()
    # keys of the 63 samples of every node: per stencil point e, the point (sign 0) then its 8 corners
    combos = [(e, c) for e in STENCIL for c in ((0, 0, 0),) + CORNER_SIGNS]
    keys = torch.empty((ng, len(combos)), dtype=torch.int64, device=device)
    for q, (e, c) in enumerate(combos):
        kx = Ix[:, e[0] + 1, c[0] + 1].to(torch.int64)
        ky = Iy[:, e[1] + 1, c[1] + 1].to(torch.int64)
        kz = Iz[:, e[2] + 1, c[2] + 1].to(torch.int64)
        keys[:, q] = ((kx[:, None, None] * My + ky[None, :, None]) * Mz + kz[None, None, :]).reshape(-1)
    uniq, inv = torch.unique(keys.reshape(-1), sorted=True, return_inverse=True)
    del keys
    pos = torch.stack([Ux[uniq // (My * Mz)], Uy[(uniq // Mz) % My], Uz[uniq % Mz]], dim=1)
    del uniq
    vals = sample(phi, pos)[inv].reshape(ng, 7, 9)
    del inv, pos
    phis = torch.empty((grid.npts, 7), dtype=torch.float64, device=device)
    phic = torch.empty((grid.npts, 7, 8), dtype=torch.float64, device=device)
    phis[:ng] = vals[:, :, 0]
    phic[:ng] = vals[:, :, 1:]
    del vals
    if grid.npts > ng:
        Se = S[ng:]
        phis[ng:] = sample(phi, Se)
        phic[ng:] = sample(phi, Se[:, :, None, :] + _offsets(CORNER_SIGNS, device)[None, None] * half)
    return phis.contiguous(), phic.contiguous()


def build_level(grid, h, bounds, cbs, batch_size, device="cuda", coeffs=False):
    """DeviceRecords of one zoom level.

    coeffs: also write the 26 cell coefficients of every row's point (the
    learned preconditioner's input, rec.row_coeffs); the centre-cell geometry is
    then computed for crossed box points too.

    grid: records.DeviceGrid (grid nodes and extra points, one row each); h: (3,) float32 stencil spacing; bounds: level-set
    grid bounds; cbs: dict of callbacks (phi, mu_m, mu_p, k_m, k_p, f_m, f_p,
    alpha, beta, g), coordinate-first."""
    import torch

    L = N.lib()
    st = N.current_stream()
    h = np.asarray(h, np.float32)
    ht = torch.tensor(h, dtype=torch.float32, device=device)
    nx, ny, nz = grid.shape
    npts = grid.npts  # grid nodes, then the extra (batch padding) points
    P = grid.points()  # (N,3) f32

    # stencil points and the corners of their cells (float32, reference order of operations)
    S = P[:, None, :] + _offsets(STENCIL, device)[None] * ht  # (N,7,3)
    half = ht * 0.5
    phis, phic = stencil_corner_phi(grid, S, ht, half, cbs["phi"])
    region = torch.empty((npts, 7), dtype=torch.int8, device=device)
    N.check(L.dips_geom_regions(N.ptr(phic), npts * 7, N.ptr(region), st), "dips_geom_regions")

    # box points (float32 test against the level-set grid bounds, discretization.py:319-333)
    b = torch.tensor(np.asarray(bounds, np.float32), device=device)
    tol = ht * np.float32(1e-6)
    bnd = ((P - b[0::2]).abs() < tol).any(dim=1) | ((P - b[1::2]).abs() < tol).any(dim=1)

    cut = (region == 0) & (~bnd)[:, None]
    cs_idx = torch.nonzero(cut.reshape(-1), as_tuple=False).reshape(-1)
    ncs = int(cs_idx.numel())
    cs_pt = (cs_idx // 7).to(torch.int32).contiguous()
    cs_e = (cs_idx % 7).to(torch.int8).contiguous()
    cs_slot = torch.full((npts * 7,), -1, dtype=torch.int32, device=device)
    cs_slot[cs_idx] = torch.arange(ncs, dtype=torch.int32, device=device)
    cc_mask = (region[:, 0] == 0) if coeffs else cut[:, 0]
    cc_pts = torch.nonzero(cc_mask, as_tuple=False).reshape(-1)
    ncc = int(cc_pts.numel())
    cc_slot = torch.full((npts,), -1, dtype=torch.int32, device=device)
    cc_slot[cc_pts] = torch.arange(ncc, dtype=torch.int32, device=device)
    ip_pts = torch.nonzero(cut.any(dim=1), as_tuple=False).reshape(-1)
    nip = int(ip_pts.numel())
    ip_slot = torch.full((npts,), -1, dtype=torch.int32, device=device)
    ip_slot[ip_pts] = torch.arange(nip, dtype=torch.int32, device=device)

    f64 = dict(dtype=torch.float64, device=device)
    one = torch.zeros(1, **f64)
    # regression at crossed stencil points
    if ncs:
        s_cs = S.reshape(-1, 3)[cs_idx]  # (ncs,3)
        cube_pos = s_cs[:, None, :] + _offsets(CUBE, device)[None] * ht
        phi_cube = sample(cbs["phi"], cube_pos).contiguous()
        mu_s = torch.stack([sample(cbs["mu_m"], s_cs), sample(cbs["mu_p"], s_cs)], dim=1).contiguous()
        cs_a = torch.empty((ncs, 2, 27), **f64)
        cs_k = torch.empty((ncs, 2), **f64)
        cs_sel = torch.empty((ncs,), dtype=torch.int8, device=device)
        r_g = torch.empty((ncs, 3), **f64)
        cs_delta = torch.empty((ncs,), **f64)
        N.check(L.dips_geom_regression(ctypes.byref(grid.c), float(h[0]), float(h[1]), float(h[2]), N.ptr(cs_pt),
                                       N.ptr(cs_e), ncs, N.ptr(phi_cube), N.ptr(mu_s), PINV_RCOND, N.ptr(cs_a),
                                       N.ptr(cs_k), N.ptr(cs_sel), N.ptr(r_g), N.ptr(cs_delta), st),
                "dips_geom_regression")
        cs_vals = torch.stack([sample(cbs["alpha"], r_g), sample(cbs["beta"], r_g), sample(cbs["mu_m"], r_g),
                               sample(cbs["mu_p"], r_g)], dim=1).contiguous()
    else:
        phi_cube = cs_a = cs_k = r_g = cs_delta = cs_vals = one
        cs_sel = torch.zeros(1, dtype=torch.int8, device=device)

    # geometry of crossed centre cells
    if ncc:
        cell_pt = cc_pts.to(torch.int32).contiguous()
        vol_m = torch.empty((ncc,), **f64)
        area_m = torch.empty((ncc, 6), **f64)
        tri = torch.empty((ncc, 10, 3, 3), **f64)
        tri_a = torch.empty((ncc, 10), **f64)
        N.check(L.dips_geom_cells(ctypes.byref(grid.c), float(h[0]), float(h[1]), float(h[2]), N.ptr(cell_pt), ncc,
                                  N.ptr(phic), N.ptr(vol_m), N.ptr(area_m), N.ptr(tri), N.ptr(tri_a), st),
                "dips_geom_cells")
        beta_tri = sample(cbs["beta"], tri).contiguous()  # (ncc,10,3)
    else:
        vol_m = area_m = tri_a = beta_tri = one

    # point and face-centre samples
    fc = P[:, None, :].repeat(1, 6, 1)
    for f, (ax, sg) in enumerate(FACES):
        fc[:, f, ax] = P[:, ax] + sg * half[ax]
    mu_face = torch.stack([sample(cbs["mu_m"], fc), sample(cbs["mu_p"], fc)], dim=2).contiguous()
    del fc
    pt_vals = torch.stack([sample(cbs[k], P) for k in ("k_m", "k_p", "f_m", "f_p", "g")], dim=1).contiguous()
    scratch = torch.empty((max(nip, 1), 250), **f64)

    nnz = torch.zeros(npts, dtype=torch.int32, device=device)
    A = N.Assemble()
    A.xs, A.ys, A.zs = grid.xs.data_ptr(), grid.ys.data_ptr(), grid.zs.data_ptr()
    A.nx, A.ny, A.nz = nx, ny, nz
    A.hx, A.hy, A.hz = float(h[0]), float(h[1]), float(h[2])
    A.bnd = (ctypes.c_float * 6)(*[float(v) for v in np.asarray(bounds, np.float32)])
    A.phis, A.region, A.cc_slot = phis.data_ptr(), region.data_ptr(), cc_slot.data_ptr()
    A.vol_m, A.area_m, A.tri_a, A.beta_tri = vol_m.data_ptr(), area_m.data_ptr(), tri_a.data_ptr(), beta_tri.data_ptr()
    A.mu_face, A.pt_vals, A.cs_slot = mu_face.data_ptr(), pt_vals.data_ptr(), cs_slot.data_ptr()
    A.cs_a, A.cs_k, A.cs_sel, A.cs_delta = cs_a.data_ptr(), cs_k.data_ptr(), cs_sel.data_ptr(), cs_delta.data_ptr()
    A.cs_vals, A.phi_cube, A.scratch, A.ip_slot = cs_vals.data_ptr(), phi_cube.data_ptr(), scratch.data_ptr(), ip_slot.data_ptr()
    A.npts = npts
    A.extra, A.nextra = (grid.extra.data_ptr() if grid.extra is not None else None), grid.nextra
    A.mode = 0
    A.nnz = nnz.data_ptr()
    N.check(L.dips_geom_assemble(ctypes.byref(A), st), "dips_geom_assemble(count)")

    # row order: contiguous batches, rows sorted by entry count (stable) inside each batch
    B = min(int(batch_size), npts)
    nb = -(-npts // B)
    pid = torch.arange(npts, dtype=torch.int64, device=device)
    key = (pid // B) * 512 + nnz.to(torch.int64)
    _, order = torch.sort(key, stable=True)
    row_of = torch.empty(npts, dtype=torch.int32, device=device)
    row_of[order] = torch.arange(npts, dtype=torch.int32, device=device)
    nnz_rows = nnz[order].to(torch.int64)
    row_ptr64 = torch.zeros(npts + 1, dtype=torch.int64, device=device)
    row_ptr64[1:] = torch.cumsum(nnz_rows, 0)
    total = int(row_ptr64[-1].item())
    if total >= 2 ** 31:
        raise ValueError("too many record entries for int32 offsets")
    ent_off = row_ptr64[:-1][row_of.to(torch.int64)].contiguous()
    ent_code = torch.empty(max(total, 1), dtype=torch.int32, device=device)
    ent_w = torch.empty(max(total, 1), dtype=torch.float32, device=device)
    row_c = torch.empty(npts, dtype=torch.float32, device=device)
    row_coeffs = torch.empty((npts, 26), dtype=torch.float32, device=device) if coeffs else None
    A.mode = 1
    A.row_coeffs = row_coeffs.data_ptr() if coeffs else None
    A.ent_off, A.row_of = ent_off.data_ptr(), row_of.data_ptr()
    A.ent_code, A.ent_w, A.row_c = ent_code.data_ptr(), ent_w.data_ptr(), row_c.data_ptr()
    N.check(L.dips_geom_assemble(ctypes.byref(A), st), "dips_geom_assemble(write)")
    torch.cuda.current_stream().synchronize()

    begins = np.arange(nb, dtype=np.int64) * B
    rows = np.minimum(begins + B, npts) - begins
    small = (nnz <= BULK_MAX_ENTRIES).to(torch.int64)
    nbulk = torch.zeros(nb, dtype=torch.int64, device=device).index_add_(0, pid // B, small).cpu().numpy()
    rec = DeviceRecords.from_device(order.to(torch.int32).contiguous(), row_ptr64.to(torch.int32).contiguous(), row_c,
                                    ent_code, ent_w, total, h, B, begins, rows, nbulk, row_coeffs=row_coeffs)
    rec.stats = dict(npts=npts, ncs=ncs, ncc=ncc, nip=nip, nnz=total, boundary=int(bnd.sum().item()))
    return rec
