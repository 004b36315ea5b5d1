"""The multiresolution hash network model type (nn/hash_network.py):
flax-layout parameter tree, flat layout shared with the device kernels (the +
network then the - network, kernel [in][out] then bias per layer, then the
hashgrid table), the C-ABI's shape tables, and unsupported options (no GPU)."""
import ctypes

import numpy as np
import pytest

LPBE = {"bound": 1.0, "pos_enc": "identity", "N_min": 16, "N_max": 2048, "F": 2, "T": 524288, "tv_scale": 0.0,
        "pos_levels": 16, "layer_widths": [16], "sol_skip_in_layers": [], "sol_act": "linear", "highest_sh_order": 2}


def test_model_tree_and_flat_layout():
    from jax_dips import _native
    from jax_dips.nn.configure import get_model
    from oracle import mlp as M

    forward, framework = get_model({"model_type": "multiresolution_hash_network",
                                    "multiresolution_hash_network": LPBE}, "multiresolution_hash_network")
    assert framework == "flax"
    model = forward.model
    assert model.dims == (1, 16, 1, 16) and model.act == 1
    tree = forward.init(7)
    assert set(tree["params"]) == {"sol_p", "sol_m"}
    assert tree["params"]["sol_p"]["Dense_0"]["kernel"].shape == (3, 16)
    assert tree["params"]["sol_m"]["Dense_1"]["kernel"].shape == (16, 1)
    flat = model.flatten(tree)
    lay = M.layout(model.dims)
    assert flat.size == lay["size"]
    # the + net's first kernel leads the flat vector, the - net's output bias ends it
    assert np.array_equal(flat[:48], tree["params"]["sol_p"]["Dense_0"]["kernel"].reshape(-1))
    assert flat[-1] == tree["params"]["sol_m"]["Dense_1"]["bias"][0]
    assert np.array_equal(model.flatten(model.unflatten(flat)), flat)
    lim = np.sqrt(6.0 / (3 + 16))
    assert np.all(np.abs(tree["params"]["sol_p"]["Dense_0"]["kernel"]) <= lim)
    net = _native.Net(*model.dims, 0, 0, 0.0, model.act)
    lib = _native.lib()
    assert lib.dips_tile_supported(ctypes.byref(net)) == 1
    assert lib.dips_net_supported(ctypes.byref(net)) == 0 and lib.dips_mfma_supported(ctypes.byref(net)) == 0
    assert lib.dips_param_count(ctypes.byref(net)) == flat.size
    assert lib.dips_param_count(ctypes.byref(_native.Net(*model.dims, 0, 0, 0.0, 2))) < 0


@pytest.mark.parametrize("change", [{"pos_enc": "spherical_harmonics", "layer_widths": [64]},
                                    {"pos_enc": "frequency", "layer_widths": [32]},
                                    {"sol_act": "sigmoid"}, {"layer_widths": [32]}, {"sol_skip_in_layers": [1]},
                                    {"pos_enc": "hashgrid", "layer_widths": [16, 16]},
                                    {"pos_enc": "hashgrid", "F": 4}, {"pos_enc": "hashgrid", "pos_levels": 17}])
def test_unsupported_options(change):
    from jax_dips.nn.hash_network import HashNetwork

    with pytest.raises(NotImplementedError):
        HashNetwork(**dict(LPBE, **change))


def test_hashgrid_tree_and_flat_layout():
    """pos_enc="hashgrid": the table under position_encoder with the reference's
    parameter name and (rows, F) shape, N_max scaled by the bound
    (model.py:124-134), the flat vector sol_p | sol_m | table, and the C-ABI's
    parameter count."""
    from jax_dips import _native
    from jax_dips.nn.configure import get_model
    from jax_dips.nn.hash_network import TABLE_KEY
    from oracle.hashgrid import HashGrid, mlp_size

    cfg = dict(LPBE, pos_enc="hashgrid", bound=1.5, T=2 ** 12, pos_levels=6, N_min=4, N_max=40)
    forward, _ = get_model({"model_type": "multiresolution_hash_network", "multiresolution_hash_network": cfg},
                           "multiresolution_hash_network")
    model = forward.model
    assert model.hashgrid == (6, 2, 2 ** 12, 4, 60, 1.5, 0) and model.in_dim == 12
    ref = HashGrid(6, 2 ** 12, 2, 4, 40, 1.5)
    tree = forward.init(3)
    tab = tree["params"]["position_encoder"][TABLE_KEY]
    assert tab.shape == (ref.rows, 2) and np.all(np.abs(tab) <= 1e-4)
    assert tree["params"]["sol_p"]["Dense_0"]["kernel"].shape == (12, 16)
    flat = model.flatten(tree)
    n_mlp = 2 * mlp_size(12, [16])
    assert flat.size == model.n_params == n_mlp + 2 * ref.rows
    assert np.array_equal(flat[n_mlp:], tab.reshape(-1))
    assert np.array_equal(model.flatten(model.unflatten(flat)), flat)
    hg = _native.HashGrid(*model.hashgrid)
    assert _native.lib().dips_hash_param_count(ctypes.byref(hg)) == flat.size
    with pytest.raises(AttributeError):
        model.dims


def test_solve_batch_rejects_hashgrid():
    from jax_dips.solvers.poisson.batch import solve_batch

    cfg = dict(LPBE, pos_enc="hashgrid", T=2 ** 10, pos_levels=4, N_min=4, N_max=32)
    model_dict = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": cfg}
    with pytest.raises(NotImplementedError):
        solve_batch([{}], num_epochs=4, model_dict=model_dict)


def test_frequency_tree_and_flat_layout():
    """pos_enc="frequency": FrequencyEncoder(L=10) (model.py:122-123) gives 60
    MLP inputs and no table; the C-ABI counts the two MLPs only."""
    from jax_dips import _native
    from jax_dips.nn.configure import get_model
    from oracle.hashgrid import mlp_size

    cfg = dict(LPBE, pos_enc="frequency")
    forward, _ = get_model({"model_type": "multiresolution_hash_network", "multiresolution_hash_network": cfg},
                           "multiresolution_hash_network")
    model = forward.model
    assert model.hashgrid[-1] == 1 and model.in_dim == 60 and model.table_rows == 0
    tree = forward.init(5)
    assert set(tree["params"]) == {"sol_p", "sol_m"}
    assert tree["params"]["sol_p"]["Dense_0"]["kernel"].shape == (60, 16)
    flat = model.flatten(tree)
    assert flat.size == model.n_params == 2 * mlp_size(60, [16])
    assert np.array_equal(model.flatten(model.unflatten(flat)), flat)
    lib = _native.lib()
    hg = _native.HashGrid(*model.hashgrid)
    assert lib.dips_hash_param_count(ctypes.byref(hg)) == flat.size
    assert lib.dips_hash_table_rows(ctypes.byref(hg)) == 0
    assert lib.dips_hash_param_count(ctypes.byref(_native.HashGrid(9, 2, 1, 1, 1, 1.0, 1))) < 0
    assert lib.dips_hash_param_count(ctypes.byref(_native.HashGrid(10, 2, 1, 1, 1, 1.0, 2))) < 0


def test_entries_ordered_by_position():
    """DeviceRecords.mfma_batches(grid): a permutation that keeps every entry in
    its (batch, phase) segment (+ entries first) and puts bit-identical
    positions fl32(p + o h) next to each other; without the grid the order is
    the phase sort alone.  Host tensors, no GPU."""
    import torch
    from jax_dips.records import DeviceGrid, DeviceRecords

    rng = np.random.default_rng(3)
    n = 6
    xs = np.linspace(-1, 1, n, dtype=np.float32)
    grid = DeviceGrid(xs, xs, xs, extra=rng.uniform(-1, 1, (5, 3)), device="cpu")
    h = np.full(3, xs[1] - xs[0], np.float32)
    nrows, B = 40, 16
    row_point = rng.permutation(grid.npts)[:nrows].astype(np.int32)
    counts = rng.integers(1, 9, nrows)
    row_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    nnz = int(row_ptr[-1])
    off = rng.integers(-1, 2, (nnz, 3))
    code = 25 * (off[:, 2] + 2) + 5 * (off[:, 1] + 2) + (off[:, 0] + 2) + 128 * rng.integers(0, 2, nnz)
    bb = np.arange(0, nrows, B)
    br = np.minimum(B, nrows - bb)
    T = lambda a, dt: torch.as_tensor(np.asarray(a), dtype=dt)
    rec = DeviceRecords.from_device(T(row_point, torch.int32), T(row_ptr, torch.int32),
                                    T(np.zeros(nrows), torch.float32), T(code, torch.int32),
                                    T(rng.standard_normal(nnz), torch.float32), nnz, h, B, bb, br, br)
    plain = rec.mfma_batches()
    mb = rec.mfma_batches(grid)
    perm = mb["t_csr"].numpy().astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(nnz))
    ent_row = np.repeat(np.arange(nrows), counts)
    assert np.array_equal(mb["t_row"].numpy(), ent_row[perm])
    assert np.array_equal(mb["t_code"].numpy(), code[perm])
    # the segments: batch, then + entries before - entries, as in the plain phase sort
    seg = lambda p: (ent_row[p] // B) * 2 + (1 - (code[p] >> 7) & 1)
    assert np.array_equal(seg(perm), seg(plain["t_csr"].numpy().astype(np.int64)))
    assert np.all(np.diff(seg(perm)) >= 0)
    for a, b in zip(plain["batches"], mb["batches"]):
        assert (a.t_count, a.n_plus) == (b.t_count, b.n_plus)
    # groups: maximal runs of one (segment, position) in this order; each batch's + groups then - groups
    # cover exactly its + entries then its - entries; without the grid there are no groups
    P = grid.points().numpy()
    pos = (P[row_point[ent_row[perm]]] + off[perm].astype(np.float32) * h).astype(np.float32)
    tg = mb["t_group"].numpy().astype(np.int64)
    assert tg[0] == 0 and tg[-1] == nnz and np.all(np.diff(tg) > 0)
    s = seg(perm)
    for g in range(len(tg) - 1):
        a, b = tg[g], tg[g + 1]
        assert np.all(s[a:b] == s[a]) and np.all(pos[a:b] == pos[a])
        if b < nnz:
            assert s[b] != s[a] or np.any(pos[b] != pos[a])
    assert mb["groups"] == len(tg) - 1 < nnz  # some positions coincide on this grid
    # inside a segment the positions run x-major, z fastest (the order of the grid points); in table order
    # z-major, x fastest (the hashgrid's dense table rows); both keep the same groups
    for a in range(nnz - 1):
        if s[a + 1] == s[a]:
            assert tuple(pos[a][[0, 1, 2]]) <= tuple(pos[a + 1][[0, 1, 2]])
    tb = rec.mfma_batches(grid, table_order=True)
    pt = tb["t_csr"].numpy().astype(np.int64)
    post = (P[row_point[ent_row[pt]]] + off[pt].astype(np.float32) * h).astype(np.float32)
    st = seg(pt)
    assert np.array_equal(st, s)
    for a in range(nnz - 1):
        if st[a + 1] == st[a]:
            assert tuple(post[a][[2, 1, 0]]) <= tuple(post[a + 1][[2, 1, 0]])
    assert tb["groups"] == mb["groups"]
    for bt in mb["batches"]:
        assert tg[bt.g_begin] == bt.t_begin
        assert tg[bt.g_begin + bt.g_plus] == bt.t_begin + bt.n_plus
        assert tg[bt.g_begin + bt.g_plus + bt.g_minus] == bt.t_begin + bt.t_count
    assert all(bt.t_group is None for bt in plain["batches"])
    # positions as the kernels compute them; equal positions of a segment are adjacent
    pts = grid.points().numpy()
    X = pts[row_point[ent_row[perm]]] + off[perm].astype(np.float32) * h
    keys = [(s,) + tuple(x.view(np.uint32)) for s, x in zip(seg(perm), X.astype(np.float32))]
    seen, last = set(), None
    for k in keys:
        if k != last:
            assert k not in seen, "a position recurs after another one"
            seen.add(k)
            last = k
    assert len(seen) < nnz  # the test data has repeated positions


@pytest.mark.parametrize("order", [1, 2, 8])
def test_spherical_harmonics_model_and_abi(order):
    """pos_enc = "spherical_harmonics" (model.py:135-136): highest_sh_order^2
    features, no table; the library's parameter count agrees; orders outside
    1..8 are rejected by the model and the library."""
    from jax_dips import _native
    from jax_dips.nn.configure import get_model
    from oracle.hashgrid import mlp_size

    cfg = dict(LPBE, pos_enc="spherical_harmonics", highest_sh_order=order)
    forward, _ = get_model({"model_type": "multiresolution_hash_network", "multiresolution_hash_network": cfg},
                           "multiresolution_hash_network")
    model = forward.model
    assert model.hashgrid[-1] == 2 and model.hashgrid[0] == order and model.in_dim == order * order
    tree = forward.init(3)
    assert set(tree["params"]) == {"sol_p", "sol_m"}
    assert tree["params"]["sol_m"]["Dense_0"]["kernel"].shape == (order * order, 16)
    flat = model.flatten(tree)
    assert flat.size == model.n_params == 2 * mlp_size(order * order, [16])
    lib = _native.lib()
    hg = _native.HashGrid(*model.hashgrid)
    assert lib.dips_hash_param_count(ctypes.byref(hg)) == flat.size
    assert lib.dips_hash_table_rows(ctypes.byref(hg)) == 0
    for bad in (0, 9):
        assert lib.dips_hash_param_count(ctypes.byref(_native.HashGrid(bad, 2, 1, 1, 1, 1.0, 2))) < 0
        with pytest.raises(NotImplementedError):
            get_model({"model_type": "multiresolution_hash_network",
                       "multiresolution_hash_network": dict(cfg, highest_sh_order=bad)}, "multiresolution_hash_network")
