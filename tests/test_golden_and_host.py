"""Oracle drift guard against the committed fixtures, and host-side logic of
the package that runs without a GPU (packing, batching, mesh, model params)."""
import os

import numpy as np

from conftest import ROOT


def test_oracle_reproduces_golden():
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    ref = np.load(os.path.join(ROOT, "tests", "golden", "sphere8_records.npz"))
    now = mg.build()
    for k in ref.files:
        assert np.allclose(ref[k], now[k], rtol=1e-12, atol=1e-14), k


def test_pack_dense_layout():
    from jax_dips.records import BULK_MAX_ENTRIES, pack_dense

    rng = np.random.default_rng(0)
    n = 50
    W = np.zeros((n, 250))
    for i in range(n):
        k = rng.choice(250, size=rng.integers(1, 30), replace=False)
        W[i, k] = rng.normal(size=k.size)
    c = rng.normal(size=n)
    pk = pack_dense(W, c, 16)
    assert pk["batch_size"] == 16 and list(pk["batch_rows"]) == [16, 16, 16, 2]
    # rows stay inside their batch and are sorted by entry count
    ptr = pk["row_ptr"]
    cnt = np.diff(ptr)
    for b0, nr, nb in zip(pk["batch_begin"], pk["batch_rows"], pk["batch_nbulk"]):
        seg = pk["row_point"][b0:b0 + nr]
        assert set(seg) == set(range(b0, b0 + nr))
        assert np.all(np.diff(cnt[b0:b0 + nr]) >= 0)
        assert nb == (cnt[b0:b0 + nr] <= BULK_MAX_ENTRIES).sum()
    # decode back to the dense weights
    back = np.zeros_like(W)
    for r in range(n):
        for e in range(ptr[r], ptr[r + 1]):
            code = int(pk["ent_code"][e])
            back[pk["row_point"][r], 2 * (code & 127) + (code >> 7)] = pk["ent_w"][e]
    assert np.allclose(back, W.astype(np.float32))
    assert np.allclose(pk["row_c"], c[pk["row_point"]].astype(np.float32))


def test_mesh_and_model_params():
    from jax_dips.domain import mesh
    from jax_dips.nn.configure import get_model

    init_fn, point_fn = mesh.construct(3)
    x = np.linspace(-1, 1, 5, dtype=np.float32)
    g = init_fn(x, x[:4], x[:3])
    assert g.R.shape == (60, 3) and g.R_xmin_boundary.shape == (12, 3) and g.R_zmax_boundary.shape == (20, 3)
    assert np.allclose(g.R[1], [x[0], x[0], x[1]])  # 'ij' meshgrid, C order
    assert g.xmin() == -1 and g.zmax() == x[2]
    md = {"model_type": "mlp", "mlp": {"hidden_layers_m": 1, "hidden_dim_m": 1, "activation_m": "jnp.tanh",
                                       "hidden_layers_p": 2, "hidden_dim_p": 10, "activation_p": "jnp.tanh"}}
    fwd, fw = get_model(md, "mlp")
    assert fw == "haiku"
    tree = fwd.init(42, x=np.zeros(3), phi_x=0.1)
    assert set(tree) == {"double_mlp/~mlp_p_fn/linear", "double_mlp/~mlp_p_fn/linear_1",
                         "double_mlp/~mlp_p_fn/linear_2", "double_mlp/~mlp_m_fn/linear",
                         "double_mlp/~mlp_m_fn/linear_1"}
    flat = fwd.model.flatten(tree)
    assert flat.size == 167
    back = fwd.model.unflatten(flat)
    for k in tree:
        assert np.array_equal(back[k]["w"], tree[k]["w"]) and np.array_equal(back[k]["b"], tree[k]["b"])
    assert np.all(tree["double_mlp/~mlp_p_fn/linear"]["b"] == 0)
    import pytest

    with pytest.raises(NotImplementedError):
        get_model({"model_type": "mlp", "mlp": dict(md["mlp"], activation_p="nn.relu")}, "mlp")


def test_data_management_and_optimizer_config():
    import pytest

    from jax_dips.data.data_management import DatasetDict, TrainData
    from jax_dips.domain import mesh
    from jax_dips.solvers.optimizers import get_optimizer

    x = np.linspace(-1, 1, 6, dtype=np.float32)
    g = mesh.construct(3)[0](x, x, x)
    dd = DatasetDict(g.R, 100)
    b = dd.get_batched_data()
    assert b.shape == (3, 100, 3) and np.array_equal(b[0], g.R[:100])
    pad = b[2, 16:]
    assert pad.min() >= -1 and pad.max() <= 1
    td = TrainData(g)
    assert np.isclose(td.alternate_res_sequentially(8, 5, g.dx, g.dy, g.dz)[0], g.dx / 4)
    with pytest.raises(ZeroDivisionError):
        td.alternate_res_sequentially(3, 0, g.dx, g.dy, g.dz)
    assert get_optimizer("custom", learning_rate=1e-3, decay_rate=0.9)["optimizer_name"] == "custom"
    with pytest.raises(ValueError):
        get_optimizer("sgd")


def test_trainer_accepts_zero_nonlinear_operators():
    """The linear problems of the reference's examples pass nonlinear operators that
    return 0 (examples/benchmark_LPBE/coefficients.py:110-117): the drop-in trainer
    takes them as linear; any other operator raises NotImplementedError, before any
    device work (the device path has no N(u) term)."""
    import pytest

    from jax_dips import problems
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    cbs = [prob[k] for k in ("initial_value_fn", "dirichlet_bc_fn", "lvl_set_fn", "mu_m_fn", "mu_p_fn", "k_m_fn",
                             "k_p_fn", "f_m_fn", "f_p_fn", "alpha_fn", "beta_fn")]
    # the check itself (the trainer past it allocates on the GPU)
    assert trainer._zero_operator(None) and trainer._zero_operator(lambda u: 0.0)
    assert trainer._zero_operator(lambda u: 0.0 * u)
    assert not trainer._zero_operator(lambda u: np.sinh(u)) and not trainer._zero_operator(lambda u: u * u)
    assert not trainer._zero_operator(lambda u: 1.0 / 0.0)
    with pytest.raises(NotImplementedError):
        trainer.setup(*cbs, lambda u: np.sinh(u), None)
    assert callable(trainer.setup(*cbs, lambda u: 0.0, lambda u: 0.0 * u))
