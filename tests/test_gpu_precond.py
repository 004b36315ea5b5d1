"""GPU parity of the learned preconditioner (P = 0.5 + s * sigmoid(MLP(26 cell
coefficients)), nn/preconditioner.py:10-35, applied as r = P * r0) in the tile
training kernel against the CPU oracle (oracle/precond.py, oracle/trainer.py).

Records and cell coefficients come from the oracle (float64) for the kernel
tests, so they isolate the tile kernel: preconditioner forward at each row's
last lane, seeds r * P * w, and the preconditioner's weight gradients on the
matrix cores.  The geometry test checks the coefficients the FP64 geometry
kernels write, and the trainer test runs the drop-in trainer with
model_dict["preconditioner"] enabled against the oracle epoch loop.

Tolerances as for the unpreconditioned kernel: float32 networks and sums
against the float64 oracle give 1e-5 relative on the loss and 1e-4 relative
(2-norm) on the gradient; the preconditioner part of the gradient is checked
on its own to 1e-4 relative so a wrong preconditioner gradient cannot hide
behind the larger network gradient.
"""
import numpy as np
import pytest

from helpers import grid_points, oracle_level
from jax_dips import problems
from oracle import discretization as D
from oracle import mlp as M
from oracle import optim as O
from oracle import precond as PC
from oracle import trainer as T


def _device(axes, rec, batch_size, with_coeffs=True):
    from jax_dips.records import DeviceGrid, DeviceRecords, pack_dense

    packed = pack_dense(rec.W, rec.c, batch_size)
    drec = DeviceRecords(packed, rec.h)
    if with_coeffs:
        drec.set_point_coeffs(rec.coeffs)
    return DeviceGrid(*axes), drec


def _params(dims, widths, seed):
    theta = M.init_params(dims, seed=seed)
    phi = PC.init_params(widths, seed=seed + 100)
    return np.concatenate([theta, phi]).astype(np.float32), theta.size


def _check(loss, grad, ol, og, nn, what):
    rel = np.linalg.norm(grad - og) / np.linalg.norm(og)
    relp = np.linalg.norm(grad[nn:] - og[nn:]) / np.linalg.norm(og[nn:])
    print(f"{what}: loss gpu {loss:.9e} oracle {ol:.9e} grad rel {rel:.2e} precond grad rel {relp:.2e}")
    assert abs(loss - ol) <= 1e-5 * abs(ol), what
    assert rel <= 1e-4, what
    assert relp <= 1e-4, what


@pytest.mark.gpu
@pytest.mark.parametrize("dims,widths,scale", [((1, 1, 2, 10), (8, 4), 1.0), ((1, 1, 2, 10), (8,), 1.5),
                                               ((2, 10, 2, 10), (8, 4), 0.75)])
def test_precond_loss_grad(dims, widths, scale):
    from jax_dips.records import Engine

    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    n = P.shape[0]
    full, nn = _params(dims, widths, 21)
    grid, drec = _device(axes, rec, n)
    pc = dict(widths=list(widths), scale=scale)
    eng = Engine(dims, full, {}, max_rows=n, max_steps=4, precond=pc)
    assert eng.tile_path and eng.P == full.size
    loss, grad = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(full.astype(np.float64), dims, rec, np.arange(n), precond=pc)
    _check(loss, grad, ol, og, nn, f"{dims} {widths}")


@pytest.mark.gpu
def test_precond_long_and_empty_rows():
    """Rows widened past 64 entries take the one-wavefront-per-row path, which
    evaluates the preconditioner on its own; emptied rows keep their constant."""
    from jax_dips.records import Engine
    from test_gpu_tiles import _synthetic_long_rows

    dims, widths = (1, 1, 2, 10), (8, 4)
    prob = problems.sphere()
    axes, P, h, bounds, cb, rec0 = oracle_level(prob, 16, 2)
    rec, _ = _synthetic_long_rows(rec0, 5)
    rec._coeffs = rec0.coeffs
    B = 1024
    full, nn = _params(dims, widths, 23)
    grid, drec = _device(axes, rec, B)
    pc = dict(widths=list(widths), scale=1.0)
    eng = Engine(dims, full, {}, max_rows=B, max_steps=4, precond=pc)
    for b, (b0, nr) in enumerate(zip(drec.batch_begin, drec.batch_rows)):
        loss, grad = eng.tile_loss_grad(grid, drec, b, 1.0 / B)
        ol, og, _ = T.loss_and_grad(full.astype(np.float64), dims, rec, np.arange(b0, b0 + nr), precond=pc)
        _check(loss, grad, ol * nr / B, og * nr / B, nn, f"batch {b}")


@pytest.mark.gpu
def test_precond_train_level_matches_oracle_steps():
    """20 tile steps (two launches each) against the oracle loop with the
    float32 optax restatement, networks and preconditioner trained jointly."""
    from jax_dips.records import Engine
    import torch

    dims, widths = (1, 1, 2, 10), (8, 4)
    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 0)
    n = P.shape[0]
    B = 1024
    full, nn = _params(dims, widths, 25)
    grid, drec = _device(axes, rec, B)
    cfg = {"optimizer_name": "custom", "learning_rate": 1e-2,
           "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    pc = dict(widths=list(widths), scale=1.0)
    eng = Engine(dims, full, cfg, max_rows=B, max_steps=64, precond=pc)
    eng.train_level(grid, drec, 5)
    torch.cuda.synchronize()
    gl = eng.loss_hist[:20].cpu().numpy().astype(np.float64)
    gp = eng.params.cpu().numpy()
    th = full.copy()
    st = O.OptState(th.size, np.float32)
    ol = []
    for e in range(5):
        for b0 in range(0, n, B):
            l, g, _ = T.loss_and_grad(th.astype(np.float64), dims, rec, np.arange(b0, b0 + B), precond=pc)
            ol.append(l)
            th = O.step(th, g.astype(np.float32), st, cfg, np.float32)
    ol = np.array(ol)
    print("gpu", gl[:4], "oracle", ol[:4])
    assert np.all(np.abs(gl - ol) <= 1e-4 * np.abs(ol))
    assert np.linalg.norm(gp - th) <= 1e-3 * np.linalg.norm(th)
    assert np.linalg.norm(gp[nn:] - th[nn:]) <= 1e-3 * np.linalg.norm(th[nn:])
    assert int(eng.count.item()) == 20


@pytest.mark.gpu
def test_precond_single_launch_equals_two_launch():
    from jax_dips.records import Engine
    import torch

    dims, widths = (1, 1, 2, 10), (8, 4)
    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, 1)
    B = 1024
    full, _ = _params(dims, widths, 27)
    grid, drec = _device(axes, rec, B)
    cfg = {"optimizer_name": "adam", "learning_rate": 1e-3}
    out = []
    for single in (True, False):
        eng = Engine(dims, full, cfg, max_rows=B, max_steps=64, precond=dict(widths=list(widths), scale=1.0))
        eng.tile_single_launch = single
        eng.train_level(grid, drec, 3)
        torch.cuda.synchronize()
        out.append((eng.loss_hist[:12].cpu().numpy(), eng.params.cpu().numpy()))
    (l1, p1), (l2, p2) = out
    assert np.array_equal(l1, l2)
    assert np.max(np.abs(p1 - p2)) <= 1e-7 * max(1.0, np.max(np.abs(p2)))


@pytest.mark.gpu
def test_precond_requires_coefficients():
    from jax_dips import _native as N
    from jax_dips.records import Engine

    dims, widths = (1, 1, 2, 10), (8, 4)
    prob = problems.sphere()
    axes, P, h, bounds, cb, rec = oracle_level(prob, 12, 0)
    n = P.shape[0]
    full, _ = _params(dims, widths, 29)
    grid, drec = _device(axes, rec, n, with_coeffs=False)
    eng = Engine(dims, full, {}, max_rows=n, max_steps=4, precond=dict(widths=list(widths), scale=1.0))
    with pytest.raises(N.DipsError):
        eng.tile_loss_grad(grid, drec, 0, 1.0 / n)


@pytest.mark.gpu
@pytest.mark.parametrize("level", [0, 2])
def test_level_coefficients_match_oracle(level):
    """build_level(coeffs=True): the 26 coefficients of every row's point from
    the FP64 geometry kernels (stored float32) equal the oracle's, box points
    included."""
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid
    import torch

    prob = problems.sphere(mu_m=2.0, mu_p=7.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    grid = DeviceGrid(*axes)
    drec = build_level(grid, T.level_spacing(h, level), bounds, problems.oracle_callbacks(prob), P.shape[0],
                       coeffs=True)
    torch.cuda.synchronize()
    rp = drec.row_point.cpu().numpy().astype(np.int64)
    got = np.empty((P.shape[0], 26))
    got[rp] = drec.row_coeffs.cpu().numpy()
    want = rec.coeffs
    assert np.all(np.isfinite(got))
    scale = np.max(np.abs(want), axis=0) + 1e-30
    err = np.max(np.abs(got - want) / scale, axis=0)
    print("max scaled error per coefficient", err.max())
    assert np.all(err <= 1e-6)


@pytest.mark.gpu
def test_trainer_with_preconditioner_matches_oracle(tmp_path):
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer
    from test_gpu_trainer import MODEL, OPT

    prob = problems.sphere()
    dims, widths = (1, 1, 2, 10), (8, 4)
    full, nn = _params(dims, widths, 31)
    model = dict(MODEL, preconditioner={"enable": True, "layer_widths": list(widths), "scaling_coeff": 1.0})
    init_mesh_fn, _ = mesh.construct(3)
    x, y, z = problems.grid_axes(16)
    g = init_mesh_fn(x, y, z)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model, init_params=full)
    new_state, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    axes, P, h, bounds = grid_points(16)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, steps, _ = T.train(P, h, bounds, cb, dims, full, 8, 2048, OPT, precond=dict(widths=list(widths), scale=1.0))
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev = tr.flatten(tr.params)
    assert dev.size == full.size and set(tr.params["preconditioner"]["params"]) == {"Dense_0", "Dense_1", "Dense_2"}
    assert np.linalg.norm(dev - th) <= 1e-4 * np.linalg.norm(th)
    assert np.linalg.norm(dev[nn:] - th[nn:]) <= 1e-4 * np.linalg.norm(th[nn:])


@pytest.mark.gpu
@pytest.mark.parametrize("level", [0, 1])
def test_precond_with_celu_networks_loss_grad(level):
    """The reference's LPBE model in full (lpbe.yaml:48-67): CELU coordinate
    networks 3 -> [16] -> 1 (hash network, identity encoder) and the tanh
    preconditioner [8, 4], on the tile kernel of shape (1, 16, 1, 16, 8, 4, CELU)."""
    from jax_dips.nn.preconditioner import Preconditioner
    from jax_dips.records import Engine
    from test_gpu_hash_network import _glorot

    dims, widths = (1, 16, 1, 16), (8, 4)
    prob = problems.sphere(mu_m=1.0, mu_p=10.0)
    axes, P, h, bounds, cb, rec = oracle_level(prob, 16, level)
    n = P.shape[0]
    _, theta = _glorot(dims, 7)
    pcm = Preconditioner(Ds=list(widths))
    full = np.concatenate([theta, pcm.flatten(pcm.init(17))]).astype(np.float32)
    nn = theta.size
    grid, drec = _device(axes, rec, n)
    pc = dict(widths=list(widths), scale=1.0)
    eng = Engine(dims, full, {}, max_rows=n, max_steps=4, precond=pc, act=1)
    assert eng.tile_path and eng.P == full.size
    loss, grad = eng.tile_loss_grad(grid, drec, 0, 1.0 / n)
    ol, og, _ = T.loss_and_grad(full.astype(np.float64), dims, rec, np.arange(n), precond=pc, act="celu")
    _check(loss, grad, ol, og, nn, f"celu {dims} {widths} level {level}")


@pytest.mark.gpu
def test_trainer_lpbe_model_with_preconditioner_matches_oracle(tmp_path):
    """The drop-in trainer with the reference's LPBE model_dict (hash network with
    the identity encoder, widths [16], preconditioner [8, 4], scaling 1.0) on the
    LPBE problem at 16^3 on [-2.5, 2.5]^3, against the oracle epoch loop with the
    same initial parameters: per-epoch losses and final parameters (networks and
    preconditioner) to 1e-4 relative."""
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer
    from test_gpu_hash_network import LPBE, OPT, _glorot

    dims, widths = (1, 16, 1, 16), (8, 4)
    axes = problems.grid_axes(16, -2.5, 2.5)
    h = float(axes[0][1] - axes[0][0])
    prob = problems.lpbe_sphere(h)
    _, theta = _glorot(dims, 3)
    full = np.concatenate([theta, PC.init_params(widths, seed=4)]).astype(np.float32)
    nn = theta.size
    model = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": LPBE,
             "preconditioner": {"enable": True, "layer_widths": list(widths), "scaling_coeff": 1.0}}
    g = mesh.construct(3)[0](*axes)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"])
    state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=8, batch_size=2048,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=OPT,
                              model_dict=model, init_params=full)
    _, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    _, P, hh, bounds = grid_points(16, -2.5, 2.5)
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    th, ole, _, _ = T.train(P, hh, bounds, cb, dims, full, 8, 2048, OPT, precond=dict(widths=list(widths), scale=1.0),
                            act="celu")
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev = tr.flatten(tr.params)
    assert dev.size == full.size and set(tr.params["params"]) == {"sol_p", "sol_m"}
    assert set(tr.params["preconditioner"]["params"]) == {"Dense_0", "Dense_1", "Dense_2"}
    assert np.linalg.norm(dev - th) <= 1e-4 * np.linalg.norm(th)
    assert np.linalg.norm(dev[nn:] - th[nn:]) <= 1e-4 * np.linalg.norm(th[nn:])


@pytest.mark.gpu
def test_lpbe_example_default_path_matches_oracle(tmp_path):
    """The reference LPBE example's default path (examples/benchmark_LPBE/main.py:149-187, lpbe.yaml:68-96):
    scattered training points sampled inside and outside the sphere, on the box boundary and on the
    interface (as its inn sampler draws them; numpy draws here), a gstate from mesh.init_gstate_3d_manually
    with the 32^3 training spacing, the zero nonlinear operators it passes, and its model_dict (identity
    hash network, CELU [16], preconditioner [8, 4]) -- against the oracle epoch loop on the same points:
    per-epoch losses and final parameters to 1e-4 relative."""
    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer
    from test_gpu_hash_network import LPBE, _glorot

    rng = np.random.default_rng(11)

    def ball(m, r0, r1):
        v = rng.normal(size=(m, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        return v * rng.uniform(r0, r1, size=(m, 1))

    face = rng.uniform(-2.5, 2.5, size=(996, 3))
    face[np.arange(996), rng.integers(0, 3, 996)] = rng.choice([-2.5, 2.5], 996)
    P = np.concatenate([ball(1000, 0.05, 0.95), ball(1000, 1.05, 2.4), face, ball(1000, 1.0, 1.0)]).astype(np.float32)
    d = np.float32(5.0 / 32)
    g_tr = mesh.init_gstate_3d_manually(P[:, 0], P[:, 1], P[:, 2], d, d, d, None, None, None, None, None, None)
    g_lvl = mesh.construct(3)[0](*problems.grid_axes(32, -2.5, 2.5))
    prob = problems.lpbe_sphere(h=float(d))
    dims, widths = (1, 16, 1, 16), (8, 4)
    _, theta = _glorot(dims, 8)
    full = np.concatenate([theta, PC.init_params(widths, seed=9)]).astype(np.float32)
    model = {"model_type": "multiresolution_hash_network", "multiresolution_hash_network": LPBE,
             "preconditioner": {"enable": True, "layer_widths": list(widths), "scaling_coeff": 1.0}}
    opt = {"optimizer_name": "custom", "learning_rate": 1e-2, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    zero = lambda u: 0.0  # noqa: E731 -- the example's nonlinear_operator_m / _p (coefficients.py:110-117)
    init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"], prob["mu_m_fn"],
                            prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"], prob["f_p_fn"],
                            prob["alpha_fn"], prob["beta_fn"], zero, zero)
    state, solve_fn = init_fn(lvl_gstate=g_lvl, tr_gstate=g_tr, eval_gstate=g_tr, num_epochs=8, batch_size=3996,
                              checkpoint_dir=str(tmp_path / "ck"), results_dir=str(tmp_path), optimizer_dict=opt,
                              model_dict=model, init_params=full)
    _, es, le = solve_fn(sim_state=state)
    tr = solve_fn.trainer
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    bounds = (-2.5, 2.5, -2.5, 2.5, -2.5, 2.5)
    th, ole, _, _ = T.train(P, np.array([d, d, d], np.float32), bounds, cb, dims, full, 8, 3996, opt,
                            precond=dict(widths=list(widths), scale=1.0), act="celu")
    print("device", le, "oracle", ole)
    assert np.all(np.abs(le - ole) <= 1e-4 * np.abs(ole))
    dev = tr.flatten(tr.params)
    assert np.linalg.norm(dev - th) <= 1e-4 * np.linalg.norm(th)
