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


@pytest.mark.parametrize("change", [{"pos_enc": "spherical_harmonics"}, {"pos_enc": "frequency", "layer_widths": [32]},
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
