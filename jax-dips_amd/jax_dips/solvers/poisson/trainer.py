"""NBM Poisson trainer with the reference's contract, running on the HIP kernels.

``setup(11 callbacks[, nonlinear_op_m, nonlinear_op_p]) -> init_fn``;
``init_fn(lvl_gstate, tr_gstate, eval_gstate, ...) -> (PoissonSimState, solve_fn)``;
``solve_fn(sim_state) -> (PoissonSimState, epoch_store, loss_epochs)``
(jax_dips/solvers/poisson/trainer.py:980-1137).

Callbacks are functions of coordinate-first positions r (shape (3, m)) written
with the ``jax_dips.xnp`` namespace (or plain numpy), i.e. the reference's
``r[0], r[1], r[2]`` callbacks with ``jnp`` replaced by ``xnp``.

What runs where (Trainer.solve):
  * per zoom level (h = h_tr * 0.5**(epoch // (E//4)), data_management.py:319-326)
    the FP64 geometry kernels build the affine residual records of every
    training point once (level_records.build_level);
  * each epoch walks the contiguous batches (no shuffling, trainer.py:539): one
    packed-tile residual/loss/gradient kernel and one reduction + optimizer
    kernel per batch (dips_tile_train_level; dips_train_level or the MFMA
    kernels for the other network shapes), losses kept on the device;
  * loss_epochs[e] = mean of the pre-update batch losses of epoch e
    (trainer.py:553-554);
  * the solution, grad u and d_n u on eval_gstate come from dips_eval
    (trainer.py:960-977).
A short last batch is padded with random points in the data's range, as the
reference pads it (its draws come from jax's PRNGKey(0), these from numpy).
The learned preconditioner (model_dict["preconditioner"], trainer.py:229-252)
trains jointly with the networks inside the tile kernel; its parameters live
under params["preconditioner"] and its initial draw comes from numpy.
model_type "multiresolution_hash_network" runs with the identity position
encoder (CELU coordinate MLPs in the tile kernel), the hashgrid encoder
(dips_hash_train_step per batch: the MLPs and the encoder's table train
together) or the frequency and spherical-harmonics encoders (the same entry
points, no table), nn/hash_network.py.  Unsupported reference options raise
NotImplementedError: algorithm 1, lbfgs, nonlinear operators other than zero, activations
other than tanh in "mlp", and the preconditioner with the hashgrid, frequency
or spherical-harmonics encoders.
"""
import logging
import os
import pickle
import signal
import time

import numpy as np

from ... import _native
from ...data import data_management
from ...data.data_management import generate_random_points_like
from ...level_records import build_level, sample
from ...nn.configure import get_model
from ...nn.preconditioner import Preconditioner
from ...records import DeviceGrid, Engine, HashEngine, capture_epoch
from ..simulation_states import PoissonSimState, PoissonSimStateFn
from ..._jaxmd_modules import dataclasses

logger = logging.getLogger(__name__)

stop_training = False


def signalHandler(signal_num, frame):
    """SIGINT stops training after the current epoch (trainer.py:68-78)."""
    global stop_training
    stop_training = True
    logger.warning("Signal %s: training will stop after the completion of the current epoch", signal_num)


DEFAULT_OPTIMIZER = {"optimizer_name": "custom", "learning_rate": 1e-3,
                     "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
DEFAULT_MODEL = {"name": None, "model_type": "mlp",
                 "mlp": {"hidden_layers_m": 1, "hidden_dim_m": 1, "activation_m": "jnp.tanh",
                         "hidden_layers_p": 2, "hidden_dim_p": 10, "activation_p": "jnp.tanh"}}


def _callbacks(sim_state_fn):
    return dict(phi=sim_state_fn.phi_fn, mu_m=sim_state_fn.mu_m_fn, mu_p=sim_state_fn.mu_p_fn,
                k_m=sim_state_fn.k_m_fn, k_p=sim_state_fn.k_p_fn, f_m=sim_state_fn.f_m_fn,
                f_p=sim_state_fn.f_p_fn, alpha=sim_state_fn.alpha_fn, beta=sim_state_fn.beta_fn,
                g=sim_state_fn.dir_bc_fn)


class Trainer:
    """Trains the two-phase network of one interface problem on the GPU."""

    # replay each zoom level's epoch from a captured HIP graph (False: launch every step from the host)
    use_graphs = True

    def __init__(self, lvl_gstate, tr_gstate, eval_gstate, sim_state, sim_state_fn, algorithm=0,
                 mgrad_over_pgrad_scalefactor=1, lvl_set_fn=None, num_epochs=1000, multi_gpu=False,
                 batch_size=131072, checkpoint_dir="./checkpoints", checkpoint_interval=2, results_dir="./",
                 loss_plot_name="solver_loss", optimizer_dict=DEFAULT_OPTIMIZER, restart=False,
                 restart_checkpoint_dir="./checkpoints", print_rate=1, model_dict=DEFAULT_MODEL, init_params=None):
        if algorithm != 0:
            raise NotImplementedError("algorithm 1 (network-based extrapolation) is not implemented in the reference either")
        if optimizer_dict.get("optimizer_name") == "lbfgs":
            raise NotImplementedError("lbfgs (jaxopt) has no device implementation")
        pc = model_dict.get("preconditioner") or {}
        self.precond = None
        if pc.get("enable", False):
            self.precond = Preconditioner(Ds=pc["layer_widths"], out_dim=1,
                                          scaling_coeff=pc.get("scaling_coeff", 1.0))
        if not (_zero_operator(sim_state_fn.nonlinear_op_m) and _zero_operator(sim_state_fn.nonlinear_op_p)):
            raise NotImplementedError("nonlinear operators other than zero have no device implementation")
        if multi_gpu:
            logger.warning("multi_gpu=True: one GPU is used; the batch walk matches the single-GPU path")
        self.lvl_gstate, self.tr_gstate, self.eval_gstate = lvl_gstate, tr_gstate, eval_gstate
        self.sim_state, self.sim_state_fn = sim_state, sim_state_fn
        self.num_epochs = int(num_epochs)
        self.batch_size = int(batch_size)
        self.checkpoint_dir, self.checkpoint_interval = checkpoint_dir, checkpoint_interval
        self.results_dir, self.loss_plot_name, self.print_rate = results_dir, loss_plot_name, max(1, int(print_rate))
        self.optimizer_dict = optimizer_dict
        self.TD = data_management.TrainData(tr_gstate, lvl_set_fn)
        self.train_points = self.TD.train_points
        self.train_dx, self.train_dy, self.train_dz = (np.float32(tr_gstate.dx), np.float32(tr_gstate.dy),
                                                       np.float32(tr_gstate.dz))
        self.forward, self.framework = get_model(model_dict, model_type=model_dict.get("model_type", "mlp"))
        self.model = self.forward.model
        if self.precond is not None and getattr(self.model, "encoder", "identity") in ("hashgrid", "frequency", "spherical_harmonics"):
            raise NotImplementedError("the preconditioner with the hashgrid / frequency encoders has no device "
                                      "implementation")
        self.cbs = _callbacks(sim_state_fn)
        self.epoch_start = 0
        self.opt_restore = None
        if restart:
            state = self.fetch_checkpoint(restart_checkpoint_dir)
            if state is None:
                raise FileNotFoundError(f"no checkpoint in {restart_checkpoint_dir}")
            self.params = state["params"]
            self.opt_restore = state["opt_state"]
            self.epoch_start = int(state["epoch"])
            self.batch_size = int(state["batch_size"])
        elif init_params is not None:
            if isinstance(init_params, dict):
                self.params = init_params
            else:
                flat = np.asarray(init_params, np.float32)
                with_pc = self.precond is not None and flat.size > self._nn_size()
                self.params = self.unflatten(flat) if with_pc else self.model.unflatten(flat)
        else:
            self.params = self.forward.init(42)
        self.params = dict(self.params)
        self.params.setdefault("preconditioner", {})
        if self.precond is not None and not self.params["preconditioner"]:
            self.params["preconditioner"] = self.precond.init(42)
        self.epoch_store = np.arange(self.epoch_start, self.num_epochs)
        self.loss_epochs = np.zeros(self.num_epochs - self.epoch_start, dtype=np.float32)

    # ------------------------------------------------------------------ checkpoints
    @staticmethod
    def fetch_checkpoint(checkpoint_dir):
        """Latest ``checkpoint_<epoch>`` by epoch number (the reference picks the
        lexicographic max, trainer.py:334, so checkpoint_100 < checkpoint_20)."""
        if checkpoint_dir is None or not os.path.exists(checkpoint_dir):
            return None
        cks = [p for p in os.listdir(checkpoint_dir) if p.startswith("checkpoint_") and p[11:].isdigit()]
        if not cks:
            return None
        best = max(cks, key=lambda p: int(p[11:]))
        with open(os.path.join(checkpoint_dir, best), "rb") as f:
            return pickle.load(f)

    @staticmethod
    def save_checkpoint(checkpoint_dir, state):
        """pickle {opt_state, params, epoch, batch_size, resolution} to
        checkpoint_dir/checkpoint_<epoch> (trainer.py:340-351)."""
        if checkpoint_dir is None:
            return None
        os.makedirs(checkpoint_dir, exist_ok=True)
        path = os.path.join(checkpoint_dir, "checkpoint_" + str(state["epoch"]))
        with open(path, "wb") as f:
            pickle.dump(state, f)
        return path

    def flatten(self, tree):
        """Flat device parameters: the two networks, then the preconditioner."""
        flat = self.model.flatten(tree)
        if self.precond is not None:
            flat = np.concatenate([flat, self.precond.flatten(tree["preconditioner"])])
        return flat

    def _nn_size(self):
        if hasattr(self.model, "n_params"):
            return self.model.n_params
        return sum(a * b + b for tag in ("p", "m") for a, b in self.model.shapes(tag))

    def unflatten(self, flat):
        flat = np.asarray(flat, np.float32)
        nn = self._nn_size()
        pre = self.precond.unflatten(flat[nn:]) if self.precond is not None else {}
        return dict(self.model.unflatten(flat[:nn]), preconditioner=pre)

    def _state(self, engine, epoch):
        return {"opt_state": {"mu": self.unflatten(engine.m.cpu().numpy()),
                              "nu": self.unflatten(engine.v.cpu().numpy()),
                              "count": int(engine.count.item())},
                "params": self.unflatten(engine.params.cpu().numpy()),
                "epoch": int(epoch), "batch_size": self.batch_size,
                "resolution": f"{self.train_dx}, {self.train_dy}, {self.train_dz}"}

    # ------------------------------------------------------------------ training
    def solve(self):
        global stop_training
        import torch

        tr = self.tr_gstate
        n = len(tr.R)
        B = min(self.batch_size, n)
        nb = -(-n // B)
        # a short last batch is padded with random points in the data's range
        # (DatasetDict.gpu_padded_batches, data_management.py:103-173); they are
        # training points like the others and sit after the grid nodes
        extra = generate_random_points_like(np.asarray(tr.R), nb * B - n) if nb * B > n else None
        if len(tr.x) * len(tr.y) * len(tr.z) == n:
            grid = DeviceGrid(tr.x, tr.y, tr.z, extra=extra)
        else:
            # scattered training points (mesh.init_gstate_3d_manually: x, y, z are the points' coordinates,
            # R their rows), e.g. the sampled sets of examples/benchmark_LPBE/main.py:149-187
            R = np.asarray(tr.R, np.float32).reshape(-1, 3)
            grid = DeviceGrid.from_points(R if extra is None else np.concatenate([R, np.asarray(extra, np.float32)]))
        lv = self.lvl_gstate
        bounds = (lv.xmin(), lv.xmax(), lv.ymin(), lv.ymax(), lv.zmin(), lv.zmax())
        E = self.num_epochs
        steps_total = (E - self.epoch_start) * nb
        # the kernels write the loss of optimizer step c at loss_hist[c], c the global step count, which a
        # restart resumes from the checkpoint's count
        count0 = int(self.opt_restore["count"]) if self.opt_restore is not None else 0
        cap = max(1, count0 + steps_total)
        if getattr(self.model, "encoder", "identity") in ("hashgrid", "frequency", "spherical_harmonics"):
            engine = HashEngine(self.model.hashgrid, self.flatten(self.params), self.optimizer_dict, max_steps=cap)
        else:
            engine = Engine(self.model.dims, self.flatten(self.params), self.optimizer_dict, max_rows=B,
                            max_steps=cap,
                            precond=self.precond.engine_config() if self.precond is not None else None,
                            act=getattr(self.model, "act", 0))
        if self.opt_restore is not None:
            engine.m.copy_(torch.as_tensor(self.flatten(self.opt_restore["mu"]), device="cuda"))
            engine.v.copy_(torch.as_tensor(self.flatten(self.opt_restore["nu"]), device="cuda"))
            engine.count.fill_(count0)
        prev = signal.signal(signal.SIGINT, signalHandler)
        stop_training = False
        t0 = time.time()
        done = self.epoch_start
        self.timing = dict(records_s=0.0, train_s=0.0)
        try:
            epoch = self.epoch_start
            while epoch < E and not stop_training:
                lvl = epoch // (E // 4)  # ZeroDivisionError for E < 4, as in the reference
                last = epoch
                while last + 1 < E and (last + 1) // (E // 4) == lvl:
                    last += 1
                h = np.array(self.TD.alternate_res_sequentially(E, epoch, tr.dx, tr.dy, tr.dz), np.float32)
                tr0 = time.time()
                with _native.trace_range(f"K1 records of zoom level {lvl}"):
                    rec = build_level(grid, h, bounds, self.cbs, B, coeffs=self.precond is not None)
                torch.cuda.synchronize()
                self.timing["records_s"] += time.time() - tr0
                tt0 = time.time()
                # the level's epoch (all its batches) is captured once in a HIP graph and replayed per epoch
                graph = None
                if self.use_graphs and isinstance(engine, Engine) and last > epoch:
                    graph = capture_epoch(engine, grid, rec)
                for e in range(epoch, last + 1):
                    if stop_training:
                        break
                    with _native.trace_range(f"epoch {e} (zoom level {lvl})"):
                        self._epoch(engine, grid, rec, graph)
                    done = e + 1
                del graph
                torch.cuda.synchronize()
                self.timing["train_s"] += time.time() - tt0
                hist = engine.loss_hist.cpu().numpy()
                for e in range(epoch, done):
                    k = count0 + (e - self.epoch_start) * nb
                    self.loss_epochs[e - self.epoch_start] = hist[k:k + nb].mean()
                    if e % self.print_rate == 0:
                        logger.info(f"epoch # {e} loss is {self.loss_epochs[e - self.epoch_start]:.6e}")
                del rec
                epoch = last + 1
        finally:
            # a handler installed outside Python (a profiler's preloaded library, an embedding host) reads back
            # as None and cannot be reinstalled from here: leave ours in place then (it only sets a flag)
            if prev is not None:
                signal.signal(signal.SIGINT, prev)
        logger.info(f"solve took {time.time() - t0} (sec)")
        self.steps = int(engine.count.item()) - count0
        self.params = self.unflatten(engine.params.cpu().numpy())
        self.save_checkpoint(self.checkpoint_dir, self._state(engine, done))
        self._plot()
        self.engine = engine
        u, gu, gnu = self.evaluate_solution_and_gradients(self.params, self.eval_gstate)
        return u, gu, gnu, self.epoch_store, self.loss_epochs

    @staticmethod
    def _epoch(engine, grid, rec, graph):
        """One epoch over the level's batches: a replay of the captured graph, or the launches themselves."""
        if graph is not None:
            graph.replay()
        else:
            engine.train_level(grid, rec, 1)

    def _plot(self):
        try:
            import matplotlib

            matplotlib.use("Agg")
            import matplotlib.pyplot as plt
        except ImportError:
            return
        os.makedirs(self.results_dir, exist_ok=True)
        fig, ax = plt.subplots(figsize=(6, 4))
        ax.plot(self.epoch_store, self.loss_epochs)
        ax.set_yscale("log")
        ax.set_xlabel("epoch")
        ax.set_ylabel("loss")
        fig.savefig(os.path.join(self.results_dir, f"{self.loss_plot_name}.png"), dpi=100)
        plt.close(fig)

    # ------------------------------------------------------------------ evaluation
    def evaluate_solution_fn(self, params, R_flat):
        from ...evaluation import evaluate_points
        import torch

        R = np.asarray(R_flat, np.float32)
        phase = sample(self.cbs["phi"], torch.as_tensor(R, device="cuda")).cpu().numpy() >= 0
        u, _ = evaluate_points(self.model, self.model.flatten(params), R, phase)
        return u

    def evaluate_solution_and_gradients(self, params, eval_gstate):
        from ...evaluation import central_normals, evaluate_points
        import torch

        R = np.asarray(eval_gstate.R, np.float32)
        phase = sample(self.cbs["phi"], torch.as_tensor(R, device="cuda")).cpu().numpy() >= 0
        nrm = central_normals(self.cbs["phi"], R, eval_gstate.dx, eval_gstate.dy, eval_gstate.dz)
        u, g, gn = evaluate_points(self.model, self.model.flatten(params), R, phase, normals=nrm)
        return u, g, gn


def _zero_operator(fn):
    """True for a missing operator or one that returns 0 wherever it is probed (the linear problems of the
    reference's examples pass such operators, e.g. examples/benchmark_LPBE/coefficients.py:110-117); its term
    N(u) V in the residual (discretization.py:369) then vanishes.  Any other operator, or one that fails on a
    float, is taken as non-zero."""
    if fn is None:
        return True
    try:
        return all(float(np.asarray(fn(u), np.float64).reshape(-1)[0]) == 0.0 for u in (0.0, 1.0, -1.0, 0.37, -25.0))
    except Exception:  # noqa: BLE001 -- any failure: not provably zero
        return False


def setup(initial_value_fn, dirichlet_bc_fn, lvl_set_fn, mu_m_fn_, mu_p_fn_, k_m_fn_, k_p_fn_, f_m_fn_, f_p_fn_,
          alpha_fn_, beta_fn_, nonlinear_op_m=None, nonlinear_op_p=None):
    """Returns init_fn (trainer.py:980-1137).  Missing nonlinear operators are zero; operators that are not
    identically zero raise NotImplementedError here, before any device work."""
    if not (_zero_operator(nonlinear_op_m) and _zero_operator(nonlinear_op_p)):
        raise NotImplementedError("nonlinear operators other than zero have no device implementation")
    if nonlinear_op_m is None:
        logger.warning("nonlinear_op_m(u) is not defined. Setting it to zero.")
    if nonlinear_op_p is None:
        logger.warning("nonlinear_op_p(u) is not defined. Setting it to zero.")
    sim_state_fn = PoissonSimStateFn(initial_value_fn, dirichlet_bc_fn, lvl_set_fn, mu_m_fn_, mu_p_fn_, k_m_fn_,
                                     k_p_fn_, f_m_fn_, f_p_fn_, alpha_fn_, beta_fn_, nonlinear_op_m, nonlinear_op_p)

    def init_fn(lvl_gstate=None, tr_gstate=None, eval_gstate=None, num_epochs=1000, batch_size=131072, algorithm=0,
                mgrad_over_pgrad_scalefactor=1, multi_gpu=False, checkpoint_interval=1000,
                checkpoint_dir="./checkpoints", results_dir="./", loss_plot_name="solver_loss",
                optimizer_dict=DEFAULT_OPTIMIZER, model_dict=DEFAULT_MODEL, restart=False,
                restart_checkpoint_dir="./checkpoints", print_rate=1, init_params=None):
        import torch

        R = torch.as_tensor(np.asarray(eval_gstate.R, np.float32), device="cuda")

        def ev(fn):
            return sample(fn, R).cpu().numpy().astype(np.float32)

        state = PoissonSimState(ev(lvl_set_fn), ev(initial_value_fn), ev(dirichlet_bc_fn), ev(mu_m_fn_),
                                ev(mu_p_fn_), ev(k_m_fn_), ev(k_p_fn_), ev(f_m_fn_), ev(f_p_fn_), ev(alpha_fn_),
                                ev(beta_fn_), None, None)

        def solve_fn(sim_state):
            trainer = Trainer(lvl_gstate, tr_gstate, eval_gstate, sim_state, sim_state_fn, algorithm,
                              mgrad_over_pgrad_scalefactor=mgrad_over_pgrad_scalefactor, lvl_set_fn=lvl_set_fn,
                              num_epochs=num_epochs, multi_gpu=multi_gpu, batch_size=batch_size,
                              checkpoint_dir=checkpoint_dir, checkpoint_interval=checkpoint_interval,
                              results_dir=results_dir, loss_plot_name=loss_plot_name, optimizer_dict=optimizer_dict,
                              restart=restart, restart_checkpoint_dir=restart_checkpoint_dir, print_rate=print_rate,
                              model_dict=model_dict, init_params=init_params)
            solve_fn.trainer = trainer
            final_solution, grad_u, grad_u_n, epoch_store, loss_epochs = trainer.solve()
            new_state = dataclasses.replace(sim_state, solution=final_solution, grad_solution=grad_u,
                                            grad_normal_solution=grad_u_n)
            return new_state, epoch_store, loss_epochs

        return state, solve_fn

    return init_fn
