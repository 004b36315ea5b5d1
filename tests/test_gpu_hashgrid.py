"""GPU parity of the hashgrid network (dips_hash_*; nn/hash_network.py with
pos_enc="hashgrid") against the CPU oracle (oracle/hashgrid.py): the loss and
gradient of one batch (MLP weights and the encoder's table), evaluation with
d u / d x, optimizer steps, and the drop-in trainer.

Tolerances: float32 device against the float64 oracle; the table gradient's
summation order is not fixed (per-bucket LDS sums, or atomic adds for tables
above 2^25 rows), so trajectories are compared to 1e-3 relative rather than
bit for bit.  Configurations: a small encoder
(4 levels, 2 dense + 2 hashed, T = 2^10) and the reference's defaults
(16 levels, T = 2^19, N_min 16, N_max 2048; model.py:96-117).
"""
import numpy as np
import pytest

from helpers import grid_points, oracle_level
from jax_dips import problems
from oracle import discretization as D
from oracle import hashgrid as HG
from oracle import trainer as T

SMALL = {"pos_levels": 4, "T": 2 ** 10, "N_min": 4, "N_max": 32}
DEFAULT = {"pos_levels": 16, "T": 2 ** 19, "N_min": 16, "N_max": 2048}
OPT = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}


def _model(cfg, seed, table_scale=1.0):
    from jax_dips.nn.hash_network import HashNetwork

    model = HashNetwork(bound=1.0, pos_enc="hashgrid", layer_widths=[16], F=2, **cfg)
    theta = model.flatten(model.init(seed))
    n_mlp = 2 * HG.mlp_size(model.in_dim, [16])
    theta[n_mlp:] *= np.float32(table_scale)
    grid = HG.HashGrid(cfg["pos_levels"], cfg["T"], 2, cfg["N_min"], cfg["N_max"], 1.0)
    assert grid.rows == model.table_rows
    return model, theta, grid, n_mlp


def _rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,level,table_scale,B,b", [(SMALL, 0, 1e4, 4096, 0), (SMALL, 2, 1e4, 1024, 2),
                                                       (SMALL, 1, 1.0, 4096, 0), (DEFAULT, 1, 1e4, 2048, 1)])
def test_loss_grad_matches_oracle(cfg, level, table_scale, B, b):
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    model, theta, grid, n_mlp = _model(cfg, 3, table_scale)
    print(f"hashgrid {cfg} level {level} table x{table_scale} batch {b} of {B}", flush=True)
    eng = HashEngine(model.hashgrid, theta, {}, max_steps=2)
    dgrid, drec = DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h)
    loss, grad = eng.loss_grad(dgrid, drec, b, 1.0 / B)
    rows = np.arange(b * B, (b + 1) * B)
    ol, og, _ = HG.loss_and_grad(theta.astype(np.float64), grid, [16], rec, rows)
    rm, rt = _rel(grad[:n_mlp], og[:n_mlp]), _rel(grad[n_mlp:], og[n_mlp:])
    print(f"  loss gpu {loss:.9e} oracle {ol:.9e}; grad rel mlp {rm:.2e} table {rt:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert rm <= 1e-4 and rt <= 1e-4
    # table rows no entry of the batch touches get exactly zero
    assert np.array_equal(grad[n_mlp:] == 0, og[n_mlp:] == 0)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,level", [(DEFAULT, 0), (SMALL, 1)])
def test_entry_order_invariance(cfg, level):
    """The loss and gradient do not depend on the entry order inside a (batch,
    phase) segment: the phase sort alone (entries at one position rarely
    adjacent) and the order by position (such entries adjacent; the kernels
    send one set of table-gradient records for each run of them) agree with
    each other and with the oracle.  Zoom level 0 of a float32 linspace grid
    puts most stencil neighbours bit-exactly on the next point."""
    import ctypes

    import torch

    from jax_dips import _native as N
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    model, theta, grid, n_mlp = _model(cfg, 5, 1e4)
    B, b = 2048, 0
    eng = HashEngine(model.hashgrid, theta, {}, max_steps=2)
    dgrid, drec = DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h)
    eng.workspace(drec, dgrid)
    out = {}
    for name, mb in (("phase", drec.mfma_batches()), ("position", drec.mfma_batches(dgrid, table_order=True)),
                     ("grid position", drec.mfma_batches(dgrid))):
        print(f"hashgrid {cfg} level {level}: {name} order", flush=True)
        g = torch.empty(eng.P, dtype=torch.float32, device="cuda")
        loss = torch.empty(1, dtype=torch.float64, device="cuda")
        N.check(N.lib().dips_hash_loss_grad(ctypes.byref(dgrid.c), ctypes.byref(drec.c), ctypes.byref(mb["batches"][b]),
                                            ctypes.byref(eng.hg), ctypes.c_float(1.0 / B), N.ptr(eng.params), N.ptr(g),
                                            N.ptr(loss), N.ptr(eng._work), int(eng._work.numel()), N.ptr(eng._dwork),
                                            int(eng._dwork.numel()), N.current_stream()), "dips_hash_loss_grad")
        torch.cuda.synchronize()
        out[name] = (float(loss.item()), g.double().cpu().numpy())
    assert not np.array_equal(drec.mfma_batches()["t_csr"].cpu().numpy(),
                              drec.mfma_batches(dgrid, table_order=True)["t_csr"].cpu().numpy())
    (l0, g0), (l1, g1) = out["phase"], out["position"]
    rows = np.arange(b * B, (b + 1) * B)
    ol, og, _ = HG.loss_and_grad(theta.astype(np.float64), grid, [16], rec, rows)
    print(f"  loss {l0:.9e} / {l1:.9e} oracle {ol:.9e}; grad rel {_rel(g1, g0):.2e}")
    assert abs(l0 - l1) <= 1e-6 * abs(ol) and abs(l1 - ol) <= 1e-5 * abs(ol)
    assert _rel(g1[:n_mlp], g0[:n_mlp]) <= 1e-5 and _rel(g1[n_mlp:], g0[n_mlp:]) <= 1e-5
    assert _rel(g1[:n_mlp], og[:n_mlp]) <= 1e-4 and _rel(g1[n_mlp:], og[n_mlp:]) <= 1e-4
    assert np.array_equal(g1[n_mlp:] == 0, og[n_mlp:] == 0)


@pytest.mark.gpu
def test_slot_overflow_matches_oracle():
    """Table-gradient records past a (workgroup, bucket) slot's capacity take
    atomic adds.  With the capacity forced down to 8 records
    (DIPS_HASH_SLOT_CAP, read once per process, hence a child test run) most
    records overflow, and the oracle comparisons of the loss / gradient tests
    still hold."""
    import os
    import subprocess
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    env = dict(os.environ, DIPS_HASH_SLOT_CAP="8")
    cmd = [sys.executable, "-u", "-m", "pytest", os.path.join(here, "test_gpu_hashgrid.py"), "-x", "-q", "-p",
           "no:cacheprovider", "-k", "test_loss_grad_matches_oracle or test_entry_order_invariance"]
    print("child run with DIPS_HASH_SLOT_CAP=8", flush=True)
    r = subprocess.run(cmd, env=env, cwd=os.path.dirname(here), capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert " passed" in r.stdout and "6 passed" in r.stdout


@pytest.mark.gpu
def test_large_table_atomic_path_matches_oracle():
    """A table above 2^25 rows (a dense level of 256^3 rows and a hashed level of
    T = 2^25 rows) takes the atomic-add table gradient instead of the buckets."""
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense

    cfg = {"pos_levels": 2, "T": 2 ** 25, "N_min": 256, "N_max": 512}
    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 12, 1)
    model, theta, grid, n_mlp = _model(cfg, 4, 1e4)
    assert grid.rows > 2 ** 25
    n = P.shape[0]
    print(f"hashgrid {cfg}: {grid.rows} table rows, {n} training points", flush=True)
    eng = HashEngine(model.hashgrid, theta, {}, max_steps=2)
    loss, grad = eng.loss_grad(DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, n), rec.h), 0, 1.0 / n)
    ol, og, _ = HG.loss_and_grad(theta.astype(np.float64), grid, [16], rec, np.arange(n))
    rm, rt = _rel(grad[:n_mlp], og[:n_mlp]), _rel(grad[n_mlp:], og[n_mlp:])
    print(f"  loss gpu {loss:.9e} oracle {ol:.9e}; grad rel mlp {rm:.2e} table {rt:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert rm <= 1e-4 and rt <= 1e-4
    nz_dev, nz_ref = np.nonzero(grad[n_mlp:])[0], np.nonzero(og[n_mlp:])[0]
    assert np.array_equal(nz_dev, nz_ref)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg,tol", [(SMALL, 1e-3), (DEFAULT, 1e-2)])
def test_eval_matches_oracle(cfg, tol):
    from jax_dips.evaluation import evaluate_points

    model, theta, grid, n_mlp = _model(cfg, 9, 1e4)
    rng = np.random.default_rng(1)
    m = 4000
    # positions on a 2^-20 lattice so that X +- 2^-14 is exact in float32; some left of the box
    X = np.round(rng.uniform(-1.1, 1.0, (m, 3)) * 2 ** 20) / 2 ** 20
    X = X.astype(np.float32)
    phase = rng.uniform(size=m) > 0.5
    u, g = evaluate_points(model, theta, X, phase, want_grad=True)
    th = theta.astype(np.float64)
    uo = HG.forward(th, grid, [16], X, phase)
    assert np.max(np.abs(u - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))
    u2, none = evaluate_points(model, theta, X, phase)
    assert none is None and np.array_equal(u2, u)
    # central differences where no level's cell changes between X - e and X + e
    hstep = np.float32(2.0 ** -14)
    for a in range(3):
        e = np.zeros(3, np.float32)
        e[a] = hstep
        rp, _ = grid.corners(X + e)
        rm, _ = grid.corners(X - e)
        same = np.all(rp == rm, axis=(0, 2))
        fd = (HG.forward(th, grid, [16], X + e, phase) - HG.forward(th, grid, [16], X - e, phase)) / (2 * hstep)
        r = _rel(g[same, a], fd[same])
        print(f"axis {a}: {same.sum()} points, grad rel {r:.2e}")
        assert same.sum() > m // 4 and r <= tol


@pytest.mark.gpu
def test_train_steps_match_oracle():
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense
    from oracle import optim as O

    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    model, theta, grid, n_mlp = _model(SMALL, 5)
    B = 1024
    eng = HashEngine(model.hashgrid, theta, OPT, max_steps=16)
    dgrid, drec = DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h)
    eng.train_level(dgrid, drec, 2)  # 2 epochs x 4 batches
    th = theta.copy()
    st = O.OptState(th.size, np.float32)
    cfg = {"optimizer_name": "custom", "learning_rate": 1e-2, "scheduler_name": "exponential", "decay_rate": 0.9}
    losses = []
    for _ in range(2):
        for b in range(4):
            loss, grad, _ = HG.loss_and_grad(th.astype(np.float64), grid, [16], rec, np.arange(b * B, (b + 1) * B))
            th = O.step(th, grad.astype(np.float32), st, cfg, np.float32)
            losses.append(loss)
    dev = eng.loss_hist[:8].cpu().numpy()
    print("device", dev, "oracle", np.array(losses))
    assert int(eng.count.item()) == 8
    assert np.all(np.abs(dev - losses) <= 1e-3 * np.abs(losses))
    p = eng.params.cpu().numpy()
    assert _rel(p[:n_mlp], th[:n_mlp]) <= 1e-3 and _rel(p[n_mlp:], th[n_mlp:]) <= 1e-3


@pytest.mark.gpu
def test_trainer_hashgrid_matches_oracle(tmp_path):
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    hg = dict(SMALL, bound=1.0, pos_enc="hashgrid", layer_widths=[16], F=2, tv_scale=0.0, sol_act="linear")
    model_dict = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": hg}
    model, theta0, grid, n_mlp = _model(SMALL, 42)
    x, y, z = problems.grid_axes(16)
    g = mesh.construct(3)[0](x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model_dict, init_params=model.unflatten(theta0))
    st, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, None, theta0, 8, 2048, OPT,
                                loss_fn=lambda t, rec, rows: HG.loss_and_grad(t.astype(np.float64), grid, [16], rec,
                                                                              rows))
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-3 * np.abs(ole))
    assert set(tr.params["params"]) == {"sol_p", "sol_m", "position_encoder"}
    dev = tr.model.flatten(tr.params)
    assert _rel(dev, th) <= 1e-3
    phase = cb("phi", P) >= 0
    uo = HG.forward(dev.astype(np.float64), grid, [16], P, phase)
    assert np.max(np.abs(st.solution - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))


FREQ = {"bound": 1.0, "pos_enc": "frequency", "layer_widths": [16], "sol_act": "linear", "tv_scale": 0.0}


def _freq_model(seed):
    from jax_dips.nn.hash_network import HashNetwork

    model = HashNetwork(**FREQ)
    return model, model.flatten(model.init(seed)), HG.Frequency(10, 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("level,B,b", [(0, 4096, 0), (2, 1024, 3)])
def test_frequency_loss_grad_matches_oracle(level, B, b):
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    model, theta, freq = _freq_model(3)
    print(f"frequency encoder level {level} batch {b} of {B}", flush=True)
    eng = HashEngine(model.hashgrid, theta, {}, max_steps=2)
    loss, grad = eng.loss_grad(DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h), b, 1.0 / B)
    ol, og, _ = HG.loss_and_grad(theta.astype(np.float64), freq, [16], rec, np.arange(b * B, (b + 1) * B))
    r = _rel(grad, og)
    print(f"  loss gpu {loss:.9e} oracle {ol:.9e}; grad rel {r:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert r <= 1e-4


@pytest.mark.gpu
def test_frequency_eval_matches_oracle():
    from jax_dips.evaluation import evaluate_points

    model, theta, freq = _freq_model(9)
    rng = np.random.default_rng(2)
    m = 2000
    X = rng.uniform(-1.0, 1.0, (m, 3)).astype(np.float32)
    phase = rng.uniform(size=m) > 0.5
    u, g = evaluate_points(model, theta, X, phase, want_grad=True)
    th = theta.astype(np.float64)
    uo = HG.forward(th, freq, [16], X, phase)
    assert np.max(np.abs(u - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))

    class Exact:
        """the same encoding without float32 rounding, for finite differences"""
        rows, F, n_features = 0, 2, 60

        def encode(self, table, Y, dt=np.float64):
            pos = (np.asarray(Y, np.float64) + 1.0) / 2.0
            c = np.exp2(np.arange(10)) * np.float64(np.float32(np.pi))
            A = c[None, None, :] * pos[:, :, None]
            return np.stack([np.sin(A), np.cos(A)], axis=-1).reshape(Y.shape[0], 60), None

    ex = Exact()
    X64 = X.astype(np.float64)
    eps = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fd = (HG.forward(th, ex, [16], X64 + e, phase) - HG.forward(th, ex, [16], X64 - e, phase)) / (2 * eps)
        r = _rel(g[:, a], fd)
        print(f"axis {a}: grad rel {r:.2e}")
        assert r <= 1e-3


@pytest.mark.gpu
def test_trainer_frequency_matches_oracle(tmp_path):
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    model_dict = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": FREQ}
    model, theta0, freq = _freq_model(42)
    x, y, z = problems.grid_axes(16)
    g = mesh.construct(3)[0](x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model_dict, init_params=model.unflatten(theta0))
    st, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, None, theta0, 8, 2048, OPT,
                                loss_fn=lambda t, rec, rows: HG.loss_and_grad(t.astype(np.float64), freq, [16], rec,
                                                                              rows))
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-3 * np.abs(ole))
    dev = tr.model.flatten(tr.params)
    assert _rel(dev, th) <= 1e-3
    phase = cb("phi", P) >= 0
    uo = HG.forward(dev.astype(np.float64), freq, [16], P, phase)
    assert np.max(np.abs(st.solution - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))


def _sh_model(seed, order):
    from jax_dips.nn.hash_network import HashNetwork

    model = HashNetwork(bound=1.0, pos_enc="spherical_harmonics", layer_widths=[16], highest_sh_order=order)
    return model, model.flatten(model.init(seed)), HG.SphericalHarmonics(order, 1.0)


@pytest.mark.gpu
@pytest.mark.parametrize("order,level,B,b", [(2, 0, 4096, 0), (8, 0, 4096, 0), (8, 2, 1024, 3), (5, 1, 2048, 1)])
def test_spherical_harmonics_loss_grad_matches_oracle(order, level, B, b):
    """The spherical-harmonics encoder (order 2: the reference's LPBE setting,
    lpbe.yaml:60; order 8: its maximum, 64 features): loss and gradient of a
    batch against the oracle."""
    from jax_dips.records import DeviceGrid, DeviceRecords, HashEngine, pack_dense

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    model, theta, sh = _sh_model(5, order)
    assert theta.size == 2 * HG.mlp_size(order * order, [16])
    print(f"spherical harmonics order {order} level {level} batch {b} of {B}", flush=True)
    eng = HashEngine(model.hashgrid, theta, {}, max_steps=2)
    loss, grad = eng.loss_grad(DeviceGrid(*axes), DeviceRecords(pack_dense(rec.W, rec.c, B), rec.h), b, 1.0 / B)
    ol, og, _ = HG.loss_and_grad(theta.astype(np.float64), sh, [16], rec, np.arange(b * B, (b + 1) * B))
    r = _rel(grad, og)
    print(f"  loss gpu {loss:.9e} oracle {ol:.9e}; grad rel {r:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol)
    assert r <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("order", [3, 8])
def test_spherical_harmonics_eval_matches_oracle(order):
    """u against the oracle network; grad u against central finite differences
    of the oracle network on float64 positions (h = 1e-6)."""
    from jax_dips.evaluation import evaluate_points

    model, theta, sh = _sh_model(11, order)
    rng = np.random.default_rng(4)
    m = 2000
    X = rng.uniform(-1.0, 1.0, (m, 3)).astype(np.float32)
    phase = rng.uniform(size=m) > 0.5
    u, g = evaluate_points(model, theta, X, phase, want_grad=True)
    th = theta.astype(np.float64)
    uo = HG.forward(th, sh, [16], X, phase)
    assert np.max(np.abs(u - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))
    ex = HG.SphericalHarmonics(order, 1.0, exact_input=True)
    X64 = X.astype(np.float64)
    eps = 1e-6
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fd = (HG.forward(th, ex, [16], X64 + e, phase) - HG.forward(th, ex, [16], X64 - e, phase)) / (2 * eps)
        r = _rel(g[:, a], fd)
        print(f"order {order} axis {a}: grad rel {r:.2e}")
        assert r <= 1e-3


@pytest.mark.gpu
def test_trainer_spherical_harmonics_matches_oracle(tmp_path):
    """The drop-in trainer with the reference's LPBE encoder settings
    (pos_enc spherical_harmonics, highest_sh_order 2, layer_widths [16]):
    per-epoch losses and parameters against the oracle loop."""
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    prob = problems.sphere()
    cfg = {"bound": 1.0, "pos_enc": "spherical_harmonics", "layer_widths": [16], "sol_act": "linear",
           "tv_scale": 0.0, "highest_sh_order": 2}
    model_dict = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": cfg}
    model, theta0, sh = _sh_model(42, 2)
    x, y, z = problems.grid_axes(16)
    g = mesh.construct(3)[0](x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model_dict, init_params=model.unflatten(theta0))
    st, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, None, theta0, 8, 2048, OPT,
                                loss_fn=lambda t, rec, rows: HG.loss_and_grad(t.astype(np.float64), sh, [16], rec,
                                                                              rows))
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-3 * np.abs(ole))
    dev = tr.model.flatten(tr.params)
    assert _rel(dev, th) <= 1e-3
    phase = cb("phi", P) >= 0
    uo = HG.forward(dev.astype(np.float64), sh, [16], P, phase)
    assert np.max(np.abs(st.solution - uo)) <= 1e-5 * max(1.0, np.max(np.abs(uo)))
