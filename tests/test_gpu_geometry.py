"""GPU parity of the FP64 geometry / regression / assembly kernels with
the CPU oracle: the records the product builds on the device equal the
oracle's (offset, phase)-keyed affine records of every training point.

Tolerance: device weights and constants are stored as float32 (relative
rounding 6e-8); the device computes volumes/areas as |det|/6 and |axb|/2 while
the oracle uses the reference's Gram-determinant forms (equal in exact
arithmetic, ~1e-8 relative apart in float64).  The bar is 1e-6 relative to the
row's largest weight for every entry and constant, and bitwise-equal region /
boundary classification (same entry counts per class).
"""
import numpy as np
import pytest

from helpers import grid_points, oracle_level
from jax_dips import problems

CASES = [("sphere", 16, 0), ("sphere", 16, 1), ("sphere", 16, 2), ("sphere", 16, 3),
         ("star", 18, 0), ("star", 18, 2), ("sphere_mu_m_big", 16, 0), ("sphere_k", 14, 1),
         ("c3_star", 16, 0), ("c3_star", 16, 1), ("c3_star", 16, 2), ("c3_star", 16, 3), ("c3_star", 22, 1),
         ("ref_sphere", 16, 0), ("ref_sphere", 16, 1), ("ref_sphere", 16, 2), ("ref_sphere", 16, 3),
         ("ref_star", 16, 0), ("ref_star", 16, 1), ("ref_star", 16, 2), ("ref_star", 16, 3)]


def make(name):
    if name == "sphere":
        return problems.sphere()
    if name == "star":
        return problems.star(r0=0.55, amp=0.15, centre=(0.0123, -0.0171, 0.0093))
    if name == "c3_star":
        # the 128^3 star workload's interface and coefficients, centred at the origin: at zoom levels >= 1
        # stencil points and at every level cell corners lie on the x = 0 and y = 0 planes and on the polar axis
        return problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0)
    if name == "ref_sphere":
        # spatially varying mu- = y^2 log(x + 2) + 4 > mu+ = exp(-z) (tests/confs/experiment_configs.py:21-186)
        return problems.ref_sphere()
    if name == "ref_star":
        # mu- = 10 (1 + 0.2 cos(2 pi (x + y)) sin(2 pi (x - y)) cos z) > mu+ = 1 (experiment_configs.py:196-412)
        return problems.ref_star()
    if name == "sphere_mu_m_big":
        return problems.sphere(radius=0.45, mu_m=20.0, mu_p=2.0)
    if name == "sphere_k":
        return problems.sphere(radius=0.6, mu_m=1.0, mu_p=3.0, k_m=2.0, k_p=0.5, centre=(0.03, -0.02, 0.05))
    raise KeyError(name)


def device_records(prob, n, level, batch):
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid
    from oracle import trainer as T

    axes, P, h, bounds = grid_points(n)
    cbs = problems.oracle_callbacks(prob)
    grid = DeviceGrid(*axes)
    rec = build_level(grid, T.level_spacing(h, level), bounds, cbs, batch)
    return grid, rec


@pytest.mark.gpu
@pytest.mark.parametrize("name,n,level", CASES)
def test_records_match_oracle(name, n, level):
    prob = make(name)
    axes, P, h, bounds, cb, orec = oracle_level(prob, n, level)
    grid, drec = device_records(prob, n, level, P.shape[0])
    W, C = drec.to_dense()
    scale = np.maximum(np.abs(orec.W).max(axis=1), 1.0)
    err = np.abs(W - orec.W).max(axis=1) / scale
    cerr = np.abs(C - orec.c) / np.maximum(np.abs(orec.c), scale)
    print(name, n, level, drec.stats, "max w err", err.max(), "max c err", cerr.max())
    assert err.max() <= 1e-6
    assert cerr.max() <= 1e-6
    # same classification: boundary rows have one entry, bulk rows 7 (or fewer with zero weights)
    nd = (np.abs(W) > 0).sum(axis=1)
    no = (np.abs(orec.W) > 0).sum(axis=1)
    assert np.array_equal(nd <= 7, no <= 7)


@pytest.mark.gpu
def test_loss_on_device_records_matches_oracle():
    from jax_dips.records import Engine
    from oracle import mlp as M
    from oracle import trainer as T

    prob = problems.sphere()
    axes, P, h, bounds, cb, orec = oracle_level(prob, 16, 1)
    grid, drec = device_records(prob, 16, 1, 1024)
    dims = (1, 1, 2, 10)
    theta = M.init_params(dims, seed=11)
    eng = Engine(dims, theta, {}, max_rows=1024, max_steps=2)
    for b0, nr in zip(drec.batch_begin, drec.batch_rows):
        loss, grad = eng.loss_grad(grid, drec, b0, nr, 1.0 / 1024)
        ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, orec, np.arange(b0, b0 + nr))
        assert abs(loss - ol) <= 1e-5 * abs(ol)
        assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
@pytest.mark.parametrize("name,level,dims", [("ref_sphere", 0, (1, 1, 2, 10)), ("ref_sphere", 2, (1, 1, 2, 10)),
                                             ("ref_star", 1, (1, 1, 2, 10)), ("ref_star", 3, (3, 64, 3, 64)),
                                             ("ref_sphere", 1, (3, 64, 3, 64)), ("c3_star", 1, (3, 64, 3, 64))])
def test_variable_mu_loss_grad_on_device_records(name, level, dims):
    """Loss and mean gradient on records the device builds for the spatially
    varying coefficients (mu at face centres, at the stencil point s and at
    r_G all differ) equal the oracle's on its own records, through the tile
    kernel (1x1/2x10) and the MFMA kernels (3x64)."""
    from jax_dips.records import Engine
    from oracle import mlp as M
    from oracle import trainer as T

    prob = make(name)
    axes, P, h, bounds, cb, orec = oracle_level(prob, 16, level)
    B = 2048
    grid, drec = device_records(prob, 16, level, B)
    theta = M.init_params(dims, seed=13)
    eng = Engine(dims, theta, {}, max_rows=B, max_steps=2)
    for b in range(len(drec.batch_begin)):
        if eng.tile_path:
            loss, grad = eng.tile_loss_grad(grid, drec, b, 1.0 / B)
        else:
            loss, grad = eng.mfma_loss_grad(grid, drec, b, 1.0 / B)
        ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, orec, np.arange(b * B, (b + 1) * B))
        print(name, level, dims, b, f"loss {loss:.9e} {ol:.9e} grad rel {np.linalg.norm(grad - og) / np.linalg.norm(og):.2e}")
        assert abs(loss - ol) <= 1e-5 * abs(ol)
        assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
@pytest.mark.parametrize("name,level", [("sphere", 0), ("star", 2)])
def test_padded_batch_points(name, level):
    """B not dividing N: the short last batch is padded with random points in
    the data's range (data_management.py:70-76, :103-173).  Their records come
    from the same kernels at off-grid positions and equal the oracle's records
    of those positions; the padded batch's loss and gradient through the tile
    kernel equal the oracle's."""
    from jax_dips.data.data_management import generate_random_points_like
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine
    from oracle import discretization as D
    from oracle import mlp as M
    from oracle import trainer as T

    prob = make(name)
    n = 16 if name == "sphere" else 18
    axes, P, h, bounds = grid_points(n)
    B = 1000
    nb = -(-len(P) // B)
    extra = generate_random_points_like(P, nb * B - len(P))
    assert len(extra) > 0 and extra.min() >= P.min() and extra.max() <= P.max()
    grid = DeviceGrid(*axes, extra=extra)
    hl = T.level_spacing(h, level)
    drec = build_level(grid, hl, bounds, problems.oracle_callbacks(prob), B)
    assert drec.nrows == nb * B and list(drec.batch_rows) == [B] * nb
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    P_all = np.concatenate([P, extra]).astype(np.float32)
    orec = T.LevelRecords(P_all, hl, bounds, cb, np.float64)
    W, C = drec.to_dense()
    scale = np.maximum(np.abs(orec.W).max(axis=1), 1.0)
    err = np.abs(W - orec.W).max(axis=1) / scale
    cerr = np.abs(C - orec.c) / np.maximum(np.abs(orec.c), scale)
    print(name, level, "padding", len(extra), "max w err", err.max(), "(padding rows", err[len(P):].max(), ")",
          "max c err", cerr.max())
    assert err.max() <= 1e-6 and cerr.max() <= 1e-6
    dims = (1, 1, 2, 10)
    theta = M.init_params(dims, seed=23)
    eng = Engine(dims, theta, {}, max_rows=B, max_steps=2)
    b = nb - 1
    loss, grad = eng.tile_loss_grad(grid, drec, b, 1.0 / B)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, orec, np.arange(b * B, (b + 1) * B))
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
@pytest.mark.parametrize("level", [1])
def test_pb_box_records_match_oracle(level):
    """The Poisson-Boltzmann-like box (union of 40 spheres, Gaussian charges,
    mu 2 / 80, k+ = kappa^2) at 16^3 on [-6, 6]^3 Angstrom: device records equal
    the oracle's, and the MFMA kernels (3x64 networks, the configuration of the
    256^3 benchmark) give the oracle's loss and gradient on them."""
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine
    from oracle import discretization as D
    from oracle import mlp as M
    from oracle import trainer as T

    prob = problems.pb_box(n_atoms=40, half_width=6.0, seed=3)
    axes, P, h, bounds = grid_points(16, -6.0, 6.0)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    hl = T.level_spacing(h, level)
    orec = T.LevelRecords(P, hl, bounds, cb, np.float64)
    grid = DeviceGrid(*axes)
    n = P.shape[0]
    drec = build_level(grid, hl, bounds, problems.oracle_callbacks(prob), n)
    W, C = drec.to_dense()
    scale = np.maximum(np.abs(orec.W).max(axis=1), 1.0)
    err = np.abs(W - orec.W).max(axis=1) / scale
    cerr = np.abs(C - orec.c) / np.maximum(np.abs(orec.c), scale)
    print("pb", level, drec.stats, "max w err", err.max(), "max c err", cerr.max())
    assert drec.stats["nip"] > 0.1 * n  # many interface points
    assert err.max() <= 1e-6 and cerr.max() <= 1e-6
    dims = (3, 64, 3, 64)
    theta = M.init_params(dims, seed=12)
    eng = Engine(dims, theta, {}, max_rows=n, max_steps=2)
    assert eng.mfma_path
    loss, grad = eng.mfma_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, orec, np.arange(n))
    print(f"loss gpu {loss:.9e} oracle {ol:.9e} grad rel {np.linalg.norm(grad - og) / np.linalg.norm(og):.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
@pytest.mark.parametrize("level", [0, 2])
def test_lpbe_sphere_records_match_oracle(level):
    """The reference's LPBE benchmark (one charged sphere, Coulomb jump alpha,
    flux jump beta from the finite-difference normal, mu 2 / 80, k+ = 80 kappa^2,
    Debye-Hueckel boundary values) at 20^3 on [-2.5, 2.5]^3: device records
    (callbacks sampled through torch on the device) equal the oracle's, and the
    MFMA kernels give the oracle's loss and gradient on them."""
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine
    from oracle import discretization as D
    from oracle import mlp as M
    from oracle import trainer as T

    axes, P, h, bounds = grid_points(20, -2.5, 2.5)
    prob = problems.lpbe_sphere(float(h[0]))
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    hl = T.level_spacing(h, level)
    orec = T.LevelRecords(P, hl, bounds, cb, np.float64)
    grid = DeviceGrid(*axes)
    n = P.shape[0]
    drec = build_level(grid, hl, bounds, problems.oracle_callbacks(prob), n)
    W, C = drec.to_dense()
    scale = np.maximum(np.abs(orec.W).max(axis=1), 1.0)
    err = np.abs(W - orec.W).max(axis=1) / scale
    cerr = np.abs(C - orec.c) / np.maximum(np.abs(orec.c), scale)
    print("lpbe", level, drec.stats, "max w err", err.max(), "max c err", cerr.max())
    assert drec.stats["nip"] > 0
    assert err.max() <= 1e-6 and cerr.max() <= 1e-6
    dims = (3, 64, 3, 64)
    theta = M.init_params(dims, seed=13)
    eng = Engine(dims, theta, {}, max_rows=n, max_steps=2)
    loss, grad = eng.mfma_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(theta.astype(np.float64), dims, orec, np.arange(n))
    print(f"loss gpu {loss:.9e} oracle {ol:.9e} grad rel {np.linalg.norm(grad - og) / np.linalg.norm(og):.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert np.linalg.norm(grad - og) <= 1e-4 * np.linalg.norm(og)


@pytest.mark.gpu
def test_nan_callback_records_match_oracle():
    """A beta callback that returns NaN at the origin (where the padding triangles of
    the interface quadrature sit) makes the interface integral NaN; the reference
    masks it with nan_to_num(rhs / diag) (discretization.py:418-419), which zeroes the
    jump contribution of every crossed cell.  Device records built with that callback
    (sampled through torch on the device) equal the oracle's, and differ from the
    records of the finite callback exactly on the crossed cells."""
    from jax_dips import xnp
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid
    from oracle import discretization as D
    from oracle import trainer as T

    prob = problems.sphere()
    fns = problems.oracle_callbacks(prob)
    bad_fns = dict(fns)
    beta = fns["beta"]
    bad_fns["beta"] = lambda r: xnp.where((r[0] == 0) & (r[1] == 0) & (r[2] == 0), float("nan"), beta(r))
    axes, P, h, bounds = grid_points(12)
    grid = DeviceGrid(*axes)
    n = P.shape[0]
    out = {}
    for name, f in (("good", fns), ("bad", bad_fns)):
        orec = T.LevelRecords(P, h, bounds, D.Callbacks(f, np.float64), np.float64)
        drec = build_level(grid, h, bounds, f, n)
        W, C = drec.to_dense()
        scale = np.maximum(np.abs(orec.W).max(axis=1), 1.0)
        err = np.abs(W - orec.W).max(axis=1) / scale
        cerr = np.abs(C - orec.c) / np.maximum(np.abs(orec.c), scale)
        print(name, "max w err", err.max(), "max c err", cerr.max(), "finite", np.isfinite(C).all())
        assert np.isfinite(C).all() and np.isfinite(W).all()
        assert err.max() <= 1e-6 and cerr.max() <= 1e-6
        out[name] = (C, orec)
    cut = (out["good"][1].slots["geo"]["region"] == 0) & ~out["good"][1].boundary
    assert cut.sum() > 10
    assert np.any(out["good"][0][cut] != out["bad"][0][cut])
    assert np.array_equal(out["good"][0][~cut], out["bad"][0][~cut])


@pytest.mark.gpu
def test_atom_field_kernel_matches_host():
    """dips_atoms_eval (cell list on the device, float64) gives the host cell
    list's phi bit for bit (same subtraction order, no contraction) and its
    charge to 1e-13 relative (exp and summation order differ), and both equal
    the all-atom brute force where phi is below cap; also through the padded
    query sizes of the record build (m not a multiple of the workgroup)."""
    import torch

    prob = problems.pb_box(n_atoms=400, half_width=10.0, seed=6)
    f = prob["field"]
    rng = np.random.default_rng(2)
    for m in (1, 255, 5003):
        x = rng.uniform(-10.5, 10.5, (3, m))
        xd = torch.as_tensor(x, device="cuda")
        phi = f.phi(xd).cpu().numpy()
        chg = f.charge(xd).cpu().numpy()
        assert np.array_equal(phi, f.phi(x)), m
        hc = f.charge(x)
        assert np.max(np.abs(chg - hc)) <= 1e-13 * max(np.max(np.abs(hc)), 1e-300), m
        d2 = ((x.T[:, None, :] - f.c[None]) ** 2).sum(axis=2)
        assert np.array_equal(phi, np.minimum((np.sqrt(d2) - f.r[None]).min(axis=1), f.cap)), m
