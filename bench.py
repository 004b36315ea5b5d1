"""Benchmark: NBM residual+grad training points/s on one MI355X.

Workload (default, --workload star): the 128^3 star interface on [-1,1]^3,
phi = r - (0.5 + 0.1 sin(5 theta) sin(3 varphi)), mu- = 1, mu+ = 100,
k- = k+ = 1, manufactured u- = exp(-|x|^2), u+ = sin x cos y sin z (f, alpha
and beta derived), two-phase tanh MLP 3x64 per phase, batch 262144 (8
batches per epoch), optimizer clip -> Adam -> exponential decay (lr 1e-3,
decay 0.9), the four zoom levels of a 2000-epoch run.  A step is one optimizer
step on one batch of training points (residual + loss + parameter gradient +
Adam), inputs resident in HBM.  Timed steps are split evenly over the four
zoom levels and walk the batches in order, as the training loop does.

After the timed steps, the drop-in trainer (setup -> init_fn -> solve_fn)
trains the same configuration for its full epoch schedule (--accuracy-epochs,
2000) end to end -- records of every zoom level, training, evaluation on the
eval grid -- and the line reports that run's wall time and the L-inf / L2
error of the trained network against the manufactured solution
(tests/test_poisson.py:277 of the reference computes the same norms).

--workload sphere: the 128^3 sphere (phi = |x| - 0.5, mu 1/10, k = 0), MLP
1x1/2x10, batch 65536.  --workload lpbe: the reference's linearized
Poisson-Boltzmann benchmark with a known answer (one charged sphere of radius 1
in [-2.5,2.5]^3, eps 2/80, examples/benchmark_LPBE), 128^3, MLP 3x64, batch
262144; its line also carries an end-to-end run of that benchmark's own model
(hash network with the identity encoder, CELU [16], learned preconditioner
[8, 4], lr 1e-2, 20000 epochs on a 32^3 grid) with its error against the known
answer.  Other workloads: sweep, pb, hashgrid (see --help).

Prints one JSON line.  --gpus N > 1 exits: this build targets one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "jax-dips_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def mlp_fwd_flops(L, H):
    return 2 * (3 * H + (L - 1) * H * H + H)


def measured_traffic(key):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    counter summaries (FETCH_SIZE and WRITE_SIZE passes calibrated on a plain
    copy, profiles/*.md lines 'traffic_bytes_per_launch <key>: <bytes>'), or None."""
    import glob
    import re

    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(f) as fh:
            for line in fh:
                m = re.match(r"\s*traffic_bytes_per_launch (.+): (\d+)\s*$", line)
                if m and m.group(1) == key:
                    return int(m.group(2))
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--n", type=int, default=0, help="training grid points per axis (0: the workload's, 128; pb 256)")
    ap.add_argument("--batch", type=int, default=0,
                    help="batch size (0: the workload's: star 262144, sphere 65536, pb 2^20)")
    ap.add_argument("--roofline-launches", type=int, default=400)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--accuracy-epochs", type=int, default=-1,
                    help="epochs of the end-to-end drop-in trainer run whose error against the manufactured "
                         "solution the line reports (-1: the workload's schedule, 2000; 0: skip)")
    ap.add_argument("--workload", default="star", choices=["sphere", "star", "sweep", "pb", "lpbe", "hashgrid"],
                    help="star: 3x64/3x64 MLP, batch 262144 (default); sphere: 1x1/2x10 MLP, batch 65536; "
                         "sweep: 64 problems of 32^3..96^3 with 2x16/2x16 MLPs, batched vs one at a time; "
                         "pb: Poisson-Boltzmann box of 2000 atoms at 256^3, 3x64/3x64 MLP, batch 2^20; "
                         "hashgrid: sphere 128^3 with the hash network (hashgrid encoder, 16 levels, T 2^19)")
    ap.add_argument("--problems", type=int, default=64, help="sweep: number of problems")
    ap.add_argument("--epochs-per-level", type=int, default=3, help="sweep: timed epochs per zoom level")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 256 MiB device-to-device copy, 5 times (for scaling HBM counters)")
    ap.add_argument("--batch-cfg", default="auto",
                    help="sweep: shape of the batched steps, 'auto' (autotuned) or 'blk_cap,grid_blocks'")
    ap.add_argument("--graph", default="on", choices=["on", "off"],
                    help="replay each zoom level's epoch from a captured HIP graph (on) or launch every step")
    ap.add_argument("--mfma-groups", default="on", choices=["on", "off"],
                    help="MFMA path: evaluate the network once per group of record entries at one position (on), "
                         "or once per entry (off)")
    ap.add_argument("--path", default="auto", choices=["auto", "tile", "lane", "mfma"],
                    help="training-step kernels: packed-tile fused step (auto for the small networks), "
                         "lane-per-row fused step, or MFMA layer tiles (auto for the wide networks)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import ctypes

    import torch

    if args.copy_calibration:
        # a copy kernel of known size: 256 MiB read and 256 MiB written per launch
        nbytes = 256 << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 5}), flush=True)
        return

    if args.workload == "sweep":
        run_sweep(args)
        return
    if args.workload == "hashgrid":
        run_hashgrid(args)
        return

    from jax_dips import _native as N
    from jax_dips import problems
    from jax_dips.level_records import build_level
    from jax_dips.records import DeviceGrid, Engine

    n, B = args.n or 128, args.batch
    lo, hi = -1.0, 1.0
    if args.workload == "pb":
        prob = problems.pb_box(n_atoms=2000, half_width=20.0, seed=0)
        lo, hi = -20.0, 20.0
        n, B = args.n or 256, B or (1 << 20)
        mlp = {"hidden_layers_m": 3, "hidden_dim_m": 64, "activation_m": "jnp.tanh",
               "hidden_layers_p": 3, "hidden_dim_p": 64, "activation_p": "jnp.tanh"}
    elif args.workload == "lpbe":
        # the reference's LPBE benchmark (one charged sphere, known answer) on [-2.5, 2.5]^3; beta's normal
        # from central differences of phi with the training spacing, as the reference computes it
        lo, hi = -2.5, 2.5
        xs = problems.grid_axes(n, lo, hi)[0]
        prob = problems.lpbe_sphere(h=float(xs[1] - xs[0]))
        B = B or 262144
        mlp = {"hidden_layers_m": 3, "hidden_dim_m": 64, "activation_m": "jnp.tanh",
               "hidden_layers_p": 3, "hidden_dim_p": 64, "activation_p": "jnp.tanh"}
    elif args.workload == "star":
        prob = problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0)
        B = B or 262144
        mlp = {"hidden_layers_m": 3, "hidden_dim_m": 64, "activation_m": "jnp.tanh",
               "hidden_layers_p": 3, "hidden_dim_p": 64, "activation_p": "jnp.tanh"}
    else:
        prob = problems.sphere(radius=0.5, mu_m=1.0, mu_p=10.0)
        B = B or 65536
        mlp = {"hidden_layers_m": 1, "hidden_dim_m": 1, "activation_m": "jnp.tanh",
               "hidden_layers_p": 2, "hidden_dim_p": 10, "activation_p": "jnp.tanh"}
    cbs = problems.oracle_callbacks(prob)
    x, y, z = problems.grid_axes(n, lo, hi)
    grid = DeviceGrid(x, y, z)
    h0 = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]], np.float32)
    bounds = (x.min(), x.max(), y.min(), y.max(), z.min(), z.max())
    from jax_dips.nn.mlp.MLP import DoubleMLP

    model = DoubleMLP(mlp=mlp)
    dims = model.dims
    theta = model.flatten(model.init(42))
    opt = {"optimizer_name": "custom", "learning_rate": 1e-3, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}

    t0 = time.time()
    levels = []
    for lvl in range(4):
        h = (h0 * np.float32(0.5) ** lvl).astype(np.float32)
        levels.append(build_level(grid, h, bounds, cbs, B))
        torch.cuda.synchronize()
        print(f"# level {lvl} records built ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    t_records = time.time() - t0
    npts = n ** 3
    nb = levels[0].batch_begin.size
    print(f"# records: {t_records:.2f} s for 4 levels; {npts} points, {nb} batches; "
          + "; ".join(f"L{i}: nnz={r.nnz} intf={r.stats['nip']}" for i, r in enumerate(levels)), file=sys.stderr,
          flush=True)

    K = max(4, args.steps - args.steps % 4)
    eng = Engine(dims, theta, opt, max_rows=B, max_steps=args.warmup + K + args.roofline_launches + 8)
    L = N.lib()
    if args.path == "auto":
        path = "tile" if eng.tile_path else ("lane" if eng.lane_path else "mfma")
    else:
        path = args.path
    use_mfma = path == "mfma"
    # MFMA path: entries at one position evaluated once (groups of the position-ordered batches)
    mgrid = grid if args.mfma_groups == "on" else None
    if use_mfma:
        for rec in levels:
            eng.mfma_workspace(rec, grid, grouped=mgrid is not None)
    if path == "tile":
        for rec in levels:
            rec.tiles(grid)

    # each level's epoch captured once in a HIP graph (tile and MFMA paths); the timed steps replay it
    eng.mfma_grouped = mgrid is not None
    graphs = {}
    if args.graph == "on" and path in ("tile", "mfma"):
        for li, rec in enumerate(levels):
            graphs[li] = eng.capture_epoch(grid, rec)

    def run_steps(k, stream):
        # k steps split over the 4 zoom levels, batches in order (cyclic): whole epochs from the level's
        # graph, the remaining steps launched directly
        per = k // 4
        for li, rec in enumerate(levels):
            steps = per
            if li in graphs:
                for _ in range(per // nb):
                    graphs[li].replay()
                steps = per % nb
            idx = np.arange(steps) % nb
            if use_mfma:
                mb = rec.mfma_batches(mgrid)["batches"]
                for b in idx:
                    N.check(L.dips_mfma_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(mb[b]),
                                                   ctypes.byref(eng.net), ctypes.byref(eng.opt),
                                                   ctypes.c_float(1.0 / B), N.ptr(eng.params), N.ptr(eng.m),
                                                   N.ptr(eng.v), N.ptr(eng.count), N.ptr(eng.loss_hist),
                                                   int(eng.loss_hist.numel()), N.ptr(eng._mwork),
                                                   int(eng._mwork.numel()), N.ptr(eng._mdwork),
                                                   int(eng._mdwork.numel()), stream), "dips_mfma_train_step")
                continue
            if path == "tile":
                tl = rec.tiles(grid)
                btb = tl["batch_tile_begin"]
                full, part = divmod(steps, nb)
                for ne, nbat in ((full, nb), (1 if part else 0, part)):
                    if ne == 0:
                        continue
                    N.check(L.dips_tile_train_level(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(tl["c"]),
                                                    N.i64_array(btb[: nbat + 1]), nbat, rec.batch_size, ne,
                                                    ctypes.byref(eng.net), ctypes.byref(eng.opt), N.ptr(eng.params),
                                                    N.ptr(eng.m), N.ptr(eng.v), N.ptr(eng.count),
                                                    N.ptr(eng.loss_hist), int(eng.loss_hist.numel()),
                                                    N.ptr(eng.partial), eng.partial_cap, None, stream),
                            "dips_tile_train_level")
                continue
            st = L.dips_train_level(ctypes.byref(grid.c), ctypes.byref(rec.c), N.i64_array(rec.batch_begin[idx]),
                                    N.i64_array(rec.batch_rows[idx]), N.i64_array(rec.batch_nbulk[idx]), steps,
                                    rec.batch_size, 1,
                                    ctypes.byref(eng.net), ctypes.byref(eng.opt), N.ptr(eng.params), N.ptr(eng.m),
                                    N.ptr(eng.v), N.ptr(eng.count), N.ptr(eng.loss_hist), int(eng.loss_hist.numel()),
                                    N.ptr(eng.partial), eng.partial_cap, N.ptr(eng.sync), stream)
            N.check(st, "dips_train_level")

    stream = N.current_stream()
    run_steps(max(4, args.warmup - args.warmup % 4), stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    run_steps(K, stream)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    value = K * B / (ms / 1e3)
    loss_tail = eng.loss_hist[: int(eng.count.item())][-4:].cpu().numpy().tolist()

    # roofline of the dominant kernel: the fused residual/loss/gradient launch.  FLOPs per launch: 3 F_fwd
    # per network evaluation of the launch -- per record entry, or per group of entries at one position
    # where the MFMA kernels evaluate a group once; the per-entry count is reported beside it
    fp_p, fp_m = mlp_fwd_flops(dims[2], dims[3]), mlp_fwd_flops(dims[0], dims[1])
    per_launch_flops, per_launch_entry_flops = [], []
    for rec in levels:
        codes = rec.ent_code[: rec.nnz].cpu().numpy()
        ptr = rec.row_ptr.cpu().numpy()
        ph = (codes >> 7) & 1
        fl = np.where(ph == 1, fp_p, fp_m) * 3.0
        cum = np.concatenate([[0.0], np.cumsum(fl)])
        for bi, (b0, nr) in enumerate(zip(rec.batch_begin, rec.batch_rows)):
            per_launch_entry_flops.append(cum[ptr[b0 + nr]] - cum[ptr[b0]])
            if use_mfma and mgrid is not None:
                bt = rec.mfma_batches(mgrid)["batches"][bi]
                per_launch_flops.append(3.0 * (bt.g_plus * fp_p + bt.g_minus * fp_m))
            else:
                per_launch_flops.append(per_launch_entry_flops[-1])
    R = max(4, args.roofline_launches - args.roofline_launches % 4)
    nblk_rl = ctypes.c_int32(0)
    torch.cuda.synchronize()
    r0, r1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    flops_sum = entry_flops_sum = 0.0
    r0.record()
    for i in range(R):
        li = (4 * i) // R
        rec = levels[li]
        b = i % nb
        flops_sum += per_launch_flops[li * nb + b]
        entry_flops_sum += per_launch_entry_flops[li * nb + b]
        if use_mfma:
            mb = rec.mfma_batches(mgrid)["batches"]
            N.check(L.dips_mfma_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(mb[b]),
                                           ctypes.byref(eng.net), ctypes.byref(eng.opt), ctypes.c_float(1.0 / B),
                                           N.ptr(eng.params), N.ptr(eng.m), N.ptr(eng.v), N.ptr(eng.count),
                                           N.ptr(eng.loss_hist), int(eng.loss_hist.numel()), N.ptr(eng._mwork),
                                           int(eng._mwork.numel()), N.ptr(eng._mdwork), int(eng._mdwork.numel()),
                                           stream), "dips_mfma_train_step")
            continue
        if path == "tile":
            # the dominant kernel alone: the tile step kernel with gradient (its
            # optimizer kernel, one workgroup, is not part of the roofline)
            tl = rec.tiles(grid)
            tb = int(tl["batch_tile_begin"][b])
            nt = int(tl["batch_tile_begin"][b + 1]) - tb - 1
            N.check(L.dips_tile_loss_grad(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(tl["c"]), tb, nt,
                                          ctypes.byref(eng.net), N.ptr(eng.params), N.ptr(eng.partial),
                                          eng.partial_cap, 1, ctypes.byref(nblk_rl), stream), "dips_tile_loss_grad")
            continue
        N.check(L.dips_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), int(rec.batch_begin[b]),
                                  int(rec.batch_rows[b]), int(rec.batch_nbulk[b]), ctypes.byref(eng.net),
                                  ctypes.byref(eng.opt),
                                  ctypes.c_float(1.0 / B), N.ptr(eng.params), N.ptr(eng.m), N.ptr(eng.v),
                                  N.ptr(eng.count), N.ptr(eng.loss_hist), int(eng.loss_hist.numel()),
                                  N.ptr(eng.partial), eng.partial_cap, N.ptr(eng.sync), stream), "dips_train_step")
    r1.record()
    torch.cuda.synchronize()
    avg_s = r0.elapsed_time(r1) / 1e3 / R
    achieved = flops_sum / R / avg_s / 1e12

    traffic = measured_traffic(f"{path} {args.workload}{n}") if args.batch == 0 else None
    # the tile step of the small networks issues its FLOPs on the FP32 VALU (tanh and deltas) with MFMA
    # only for the weight-gradient blocks; the MFMA step of the wide networks is bound by the matrix cores
    bound = "mfma" if path == "mfma" else "valu"
    accuracy = None
    # the workload's schedule (2000 epochs); PB: 200 epochs (32 optimizer steps per epoch at 5 ms each)
    acc_epochs = (200 if args.workload == "pb" else 2000) if args.accuracy_epochs < 0 else args.accuracy_epochs
    if acc_epochs > 0:
        del levels, eng
        torch.cuda.empty_cache()
        accuracy = accuracy_run(prob, n, lo, hi, mlp, B, acc_epochs, opt)
    reference_model = None
    if args.workload == "lpbe" and acc_epochs > 0:
        # the reference's own LPBE benchmark model in full (examples/benchmark_LPBE/conf/lpbe.yaml:17-67): the
        # hash network with the identity encoder (CELU 3 -> [16] -> 1 per phase) and the learned
        # preconditioner [8, 4], learning rate 1e-2 with exponential decay, 20000 epochs, on its uniform 32^3
        # training grid (lpbe.yaml:78-81) as one batch of 32^3 points
        xs32 = problems.grid_axes(32, lo, hi)[0]
        ref_model = {"model_type": "multiresolution_hash_network",
                     "multiresolution_hash_network": {
                         "bound": 1.0, "pos_enc": "identity", "N_min": 16, "N_max": 2048, "F": 2, "T": 524288,
                         "tv_scale": 0.0, "pos_levels": 16, "layer_widths": [16], "sol_skip_in_layers": [],
                         "sol_act": "linear", "highest_sh_order": 2},
                     "preconditioner": {"enable": True, "layer_widths": [8, 4], "scaling_coeff": 1.0}}
        ref_opt = {"optimizer_name": "custom", "learning_rate": 1e-2,
                   "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
        reference_model = accuracy_run(problems.lpbe_sphere(h=float(xs32[1] - xs32[0])), 32, lo, hi, None,
                                       32 ** 3, 20000 if args.accuracy_epochs < 0 else args.accuracy_epochs,
                                       ref_opt, model_dict=ref_model)
        reference_model["model"] = ("the reference's LPBE model: hash network, identity encoder, CELU [16] per "
                                    "phase, preconditioner [8, 4]; 32^3 grid, one batch of 32768, lr 1e-2")
    result = {
        "metric": "NBM residual+grad points/s vs JAX ref on same GPU at equal Linf error",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} {n}^3 on [{lo:g},{hi:g}]^3, MLP {dims[0]}x{dims[1]}/{dims[2]}x{dims[3]} tanh, batch {B}, "
                               f"4 zoom levels (2000-epoch schedule), {path} kernels",
                   "global_batch": B, "grid": n, "parallelism": "single GPU"},
        "roofline": {"bound": bound, "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "kernel": {"mfma": "mfma step: mfma_fwd_kernel + mfma_rowsum_kernel + mfma_seed_kernel + "
                                        "mfma_bwd_kernel + mfma_reduce_kernel + mfma_apply_kernel (one optimizer "
                                        "step; forward hidden layers as exact 3-way bf16 splits on the bf16 "
                                        "matrix cores, 6 products, fp32 accumulate; backward on the FP32 matrix "
                                        "cores from the forward's stored activations; peak = the FP32 matrix "
                                        "peak)",
                                "lane": "loss_grad_kernel<grad,fused> (one launch per optimizer step)",
                                "tile": "tile_step_kernel<grad> (gradient launch of a step; the optimizer "
                                        "runs after it in a one-workgroup tile_apply_kernel); FP32 VALU "
                                        "bound (non-packed VALU ceiling 78.6 TFLOP/s)"}[path],
                     "avg_launch_us": avg_s * 1e6,
                     "flop_per_launch": flops_sum / R,
                     "entry_flop_per_launch": entry_flops_sum / R,
                     "entry_rate_tflops": entry_flops_sum / R / avg_s / 1e12},
        "records_build_s": t_records,
        "loss_tail": loss_tail,
    }
    if accuracy is not None:
        result["accuracy"] = accuracy
    if reference_model is not None:
        result["accuracy_reference_model"] = reference_model
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args.workload, n, args.cpu_seconds)
    print(json.dumps(result), flush=True)


def accuracy_run(prob, n, lo, hi, mlp, B, epochs, opt, model_dict=None):
    """The workload trained end to end through the drop-in trainer
    (setup -> init_fn -> solve_fn, initial parameters from the model's own
    initialiser, seed 42) for ``epochs`` epochs with the zoom schedule, the
    training grid as eval grid; model_dict defaults to the two-phase MLP mlp.  Returns the wall time of solve_fn (records of
    the four zoom levels, training, evaluation), its parts, the training
    throughput and the error of the trained network against the manufactured
    solution at the eval nodes: L-inf = max |u - u*|, RMS, L2 = sqrt(sum (u - u*)^2)
    and relative L2 = L2 / sqrt(sum u*^2) (the norms the reference logs,
    tests/test_poisson.py:277-281)."""
    import shutil
    import tempfile

    import torch

    from jax_dips.domain import mesh
    from jax_dips.solvers.poisson import trainer

    init_mesh_fn, _ = mesh.construct(3)
    x, y, z = problems_axes(n, lo, hi)
    g = init_mesh_fn(x, y, z)
    work = tempfile.mkdtemp(prefix="dips_bench_")
    try:
        init_fn = trainer.setup(prob["initial_value_fn"], prob["dirichlet_bc_fn"], prob["lvl_set_fn"],
                                prob["mu_m_fn"], prob["mu_p_fn"], prob["k_m_fn"], prob["k_p_fn"], prob["f_m_fn"],
                                prob["f_p_fn"], prob["alpha_fn"], prob["beta_fn"])
        state, solve_fn = init_fn(lvl_gstate=g, tr_gstate=g, eval_gstate=g, num_epochs=epochs, batch_size=B,
                                  checkpoint_dir=os.path.join(work, "ck"), results_dir=work, optimizer_dict=opt,
                                  model_dict=model_dict or {"model_type": "mlp", "mlp": mlp}, print_rate=epochs)
        torch.cuda.synchronize()
        t0 = time.time()
        new_state, _, loss_epochs = solve_fn(sim_state=state)
        torch.cuda.synchronize()
        wall = time.time() - t0
        tr = solve_fn.trainer
    finally:
        shutil.rmtree(work, ignore_errors=True)
    R = np.asarray(g.R, np.float64)
    npts = R.shape[0]
    if "exact_fn" not in prob:
        # no analytic solution (the PB box), and the unpreconditioned BiCGSTAB judge does not converge on its
        # operator (cut cells with mu 2 / 80 give rows far from diagonal dominance and negative eigenvalues):
        # the end-to-end run is reported without an error figure
        print(f"# accuracy run: {epochs} epochs in {wall:.1f} s (records {tr.timing['records_s']:.1f} s, training "
              f"{tr.timing['train_s']:.1f} s); no reference solution", file=sys.stderr, flush=True)
        return {"epochs": epochs, "steps": int(tr.steps), "end_to_end_s": wall,
                "records_s": tr.timing["records_s"], "train_s": tr.timing["train_s"],
                "train_points_per_s": npts * epochs / tr.timing["train_s"],
                "loss_first": float(loss_epochs[0]), "loss_last": float(loss_epochs[-1]),
                "reference": "none: no analytic solution, and the grid solve does not converge on this operator"}
    exact = np.asarray(prob["exact_fn"](R.T), np.float64)
    err = np.asarray(new_state.solution, np.float64) - exact
    # the trained network on the (2n)^3 eval grid (the reference's tests evaluate 128^3 runs at 256^3,
    # confs/poisson.yaml:13-15): u at nodes between the training nodes too
    x2, y2, z2 = problems_axes(2 * n, lo, hi)
    g2 = init_mesh_fn(x2, y2, z2)
    torch.cuda.synchronize()
    t1 = time.time()
    u2, _, _ = tr.evaluate_solution_and_gradients(tr.params, g2)
    torch.cuda.synchronize()
    eval_s = time.time() - t1
    R2 = np.asarray(g2.R, np.float64)
    exact2 = np.asarray(prob["exact_fn"](R2.T), np.float64)
    err2 = np.asarray(u2, np.float64) - exact2
    del R2, u2
    # the judge: the matrix-free BiCGSTAB solve of the same cut-cell equations on the grid values (the method
    # without the network), against the same exact solution at the training nodes
    from jax_dips.judge import GridJudge

    torch.cuda.synchronize()
    t2 = time.time()
    ug, iters, relres = GridJudge(g, g, problems_callbacks(prob)).solve(rtol=1e-6, max_iter=5000)
    torch.cuda.synchronize()
    judge_s = time.time() - t2
    errg = np.asarray(ug, np.float64) - exact
    judge = {"solver": "BiCGSTAB on grid values (the same discretisation, spacing h)", "iterations": int(iters),
             "relres": float(relres), "converged": bool(relres <= 1e-5), "solve_s": judge_s,
             "linf": float(np.abs(errg).max()), "rms": float(np.sqrt(np.mean(errg ** 2))),
             "rel_l2": float(np.sqrt(np.sum(errg ** 2) / np.sum(exact ** 2)))}
    print(f"# accuracy run: {epochs} epochs in {wall:.1f} s (records {tr.timing['records_s']:.1f} s, training "
          f"{tr.timing['train_s']:.1f} s); L-inf {np.abs(err).max():.4e} at the training nodes, "
          f"{np.abs(err2).max():.4e} on the {2 * n}^3 eval grid ({eval_s:.2f} s); grid solve L-inf "
          f"{judge['linf']:.4e} ({iters} iterations, relres {relres:.1e})", file=sys.stderr, flush=True)
    return {"epochs": epochs, "steps": int(tr.steps), "end_to_end_s": wall,
            "records_s": tr.timing["records_s"], "train_s": tr.timing["train_s"],
            "train_points_per_s": npts * epochs / tr.timing["train_s"],
            "linf": float(np.abs(err).max()), "rms": float(np.sqrt(np.mean(err ** 2))),
            "l2": float(np.sqrt(np.sum(err ** 2))), "rel_l2": float(np.sqrt(np.sum(err ** 2) / np.sum(exact ** 2))),
            "loss_first": float(loss_epochs[0]), "loss_last": float(loss_epochs[-1]),
            "eval_grid": f"{n}^3 training nodes",
            "eval_2n": {"grid": f"{2 * n}^3 on the same box", "linf": float(np.abs(err2).max()),
                        "rms": float(np.sqrt(np.mean(err2 ** 2))),
                        "rel_l2": float(np.sqrt(np.sum(err2 ** 2) / np.sum(exact2 ** 2))),
                        "eval_s": eval_s, "points": int(err2.size)},
            "judge": judge}


def problems_axes(n, lo, hi):
    from jax_dips import problems

    return problems.grid_axes(n, lo, hi)


def problems_callbacks(prob):
    from jax_dips import problems

    return problems.oracle_callbacks(prob)


def run_sweep(args):
    """The batched sweep: problems of the sweep workload (problems.sweep_problem,
    2x16/2x16 tanh MLPs, batch min(N, 65536), padding as the trainer pads),
    trained level by level over the four zoom levels of a 1000-epoch run,
    ``epochs_per_level`` epochs each; the same epochs run once batched
    (dips_batch_train: one step launch over all problems' current batches, one
    optimizer workgroup per problem) and once problem by problem
    (dips_tile_train_level), both from the same parameters, HIP events around
    each.  value = training points per second of the batched run (sum over
    problems of steps x batch size / time)."""
    import torch

    from jax_dips import problems
    from jax_dips.data.data_management import generate_random_points_like
    from jax_dips.level_records import build_level
    from jax_dips.nn.mlp.MLP import DoubleMLP
    from jax_dips.records import BatchEngine, DeviceGrid, Engine

    mlp = {"hidden_layers_m": 2, "hidden_dim_m": 16, "activation_m": "jnp.tanh",
           "hidden_layers_p": 2, "hidden_dim_p": 16, "activation_p": "jnp.tanh"}
    model = DoubleMLP(mlp=mlp)
    dims = model.dims
    opt = {"optimizer_name": "custom", "learning_rate": 1e-3, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    specs = [problems.sweep_problem(i) for i in range(args.problems)]
    grids, info = [], []
    for prob, n in specs:
        x, y, z = problems.grid_axes(n)
        N_ = n ** 3
        B = min(N_, args.batch or 65536)
        nb = -(-N_ // B)
        P = np.stack(np.meshgrid(x, y, z, indexing="ij"), -1).reshape(-1, 3)
        extra = generate_random_points_like(P, nb * B - N_) if nb * B > N_ else None
        grids.append(DeviceGrid(x, y, z, extra=extra))
        h0 = np.array([x[1] - x[0]] * 3, np.float32)
        info.append((problems.oracle_callbacks(prob), h0, (x.min(), x.max()) * 3, B, nb, n))
    thetas = [model.flatten(model.init(42 + i)) for i in range(len(specs))]
    E = args.epochs_per_level
    cap = [4 * (E + 1) * nb + 8 for (_, _, _, _, nb, _) in info]
    beng = BatchEngine(dims, thetas, opt, cap)
    sengs = [Engine(dims, th, opt, max_rows=1, max_steps=c) for th, c in zip(thetas, cap)]
    t_rec = t_batch = t_seq = 0.0
    points = flops = 0.0
    steps_global = 0
    for lvl in range(4):
        t0 = time.time()
        recs = []
        for g, (cbs, h0, bnd, B, nb, n) in zip(grids, info):
            recs.append(build_level(g, (h0 * np.float32(0.5) ** lvl).astype(np.float32), bnd, cbs, B))
        for g, r in zip(grids, recs):
            r.tiles(g)
        torch.cuda.synchronize()
        t_rec += time.time() - t0
        if lvl == 0:  # warm up both paths (module load, LDS attributes), then choose the batched shape
            beng.train_level(grids, recs, 1)
            for e, g, r in zip(sengs, grids, recs):
                e.train_level(g, r, 1)
            torch.cuda.synchronize()
            ta = time.time()
            cfg = beng.autotune(grids, recs) if args.batch_cfg == "auto" else tuple(
                int(v) for v in args.batch_cfg.split(","))
            torch.cuda.synchronize()
            t_tune = time.time() - ta
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        beng.train_level(grids, recs, E, cfg=cfg)
        ev[1].record()
        ev[2].record()
        for e, g, r in zip(sengs, grids, recs):
            e.train_level(g, r, E)
        ev[3].record()
        torch.cuda.synchronize()
        t_batch += ev[0].elapsed_time(ev[1]) / 1e3
        t_seq += ev[2].elapsed_time(ev[3]) / 1e3
        steps_global += E * max(nb for (_, _, _, _, nb, _) in info)
        for r, (_, _, _, B, nb, _) in zip(recs, info):
            points += E * nb * B
            codes = r.ent_code[: r.nnz].cpu().numpy()
            flops += E * 3.0 * mlp_fwd_flops(2, 16) * codes.size  # both phases are 2x16
        del recs
    value = points / t_batch
    achieved = flops / t_batch / 1e12
    result = {
        "metric": "batched sweep training points/s (64 problems, 2x16/2x16 MLP, batch min(N, 65536))",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": steps_global,
        "warmup": 1,
        "ms_per_step": t_batch * 1e3 / steps_global,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"sweep of {len(specs)} problems (sizes "
                               + ",".join(f"{k}^3 x{sum(1 for _, n in specs if n == k)}" for k in problems.SWEEP_SIZES)
                               + f"), {E} epochs per zoom level x 4 levels, tile kernels",
                   "global_batch": int(sum(B for (_, _, _, B, _, _) in info)), "parallelism": "single GPU"},
        "sequential_value": points / t_seq,
        "speedup_vs_sequential": t_seq / t_batch,
        "batch_cfg": {"blk_cap": cfg[0], "grid_blocks": cfg[1], "autotune_s": t_tune,
                      "candidates_ms_per_step": getattr(beng, "tuning", {}).get("candidates_ms")},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "tile_batch_step_kernel (persistent, work queue) + tile_batch_apply_kernel + counter "
                               "update, per global step (whole batched step; FP32 VALU bound)"},
        "records_build_s": t_rec,
        "loss_tail": [float(e.loss_hist[int(e.count.item()) - 1]) for e in beng.engines[:4]],
    }
    print(json.dumps(result), flush=True)


def run_hashgrid(args):
    """The hash network with the hashgrid encoder at the reference's defaults
    (16 levels of 2 features, T = 2^19, N_min 16, N_max 2048, one CELU layer of
    16; nn/hash_encoding/model.py:96-117) on the sphere problem at n^3, batch
    65536, the four zoom levels of a 2000-epoch run.  A step is one
    dips_hash_train_step: forward, row sums, backward (MLP gradients on the
    matrix cores, table gradient by atomic adds), reduction and Adam over every
    parameter, the table's 16.8 M entries included.

    Roofline (HBM): the step's compulsory bytes, P x 48 B (parameters, m and v
    read and written: 24; table gradient zeroed 4, scattered into once 8 and
    read 4; gradient written 4 and read 4) + 24 B per entry (phase-sorted entry
    arrays and u), over the step's time; table gathers that stay in the caches
    are not counted."""
    import ctypes

    import torch

    from jax_dips import _native as N
    from jax_dips import problems
    from jax_dips.level_records import build_level
    from jax_dips.nn.hash_network import HashNetwork
    from jax_dips.records import DeviceGrid, HashEngine

    n, B = args.n or 128, args.batch or 65536
    prob = problems.sphere(radius=0.5, mu_m=1.0, mu_p=10.0)
    cbs = problems.oracle_callbacks(prob)
    x, y, z = problems.grid_axes(n)
    grid = DeviceGrid(x, y, z)
    h0 = np.array([x[1] - x[0]] * 3, np.float32)
    bounds = (x.min(), x.max(), y.min(), y.max(), z.min(), z.max())
    model = HashNetwork(bound=1.0, pos_enc="hashgrid", layer_widths=[16], pos_levels=16, T=2 ** 19, N_min=16,
                        N_max=2048, F=2)
    theta = model.flatten(model.init(42))
    opt = {"optimizer_name": "custom", "learning_rate": 1e-3, "sched": {"scheduler_name": "exponential", "decay_rate": 0.9}}
    t0 = time.time()
    levels = []
    for lvl in range(4):
        levels.append(build_level(grid, (h0 * np.float32(0.5) ** lvl).astype(np.float32), bounds, cbs, B))
        torch.cuda.synchronize()
        print(f"# level {lvl} records built ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    t_records = time.time() - t0
    nb = levels[0].batch_begin.size
    K = max(4, args.steps - args.steps % 4)
    W = max(4, args.warmup - args.warmup % 4)
    eng = HashEngine(model.hashgrid, theta, opt, max_steps=W + K + 8)
    for rec in levels:
        eng.workspace(rec, grid)
    L = N.lib()
    stream = N.current_stream()
    inv_b = ctypes.c_float(1.0 / B)

    def run_steps(k):
        per = k // 4
        ents = 0
        for rec in levels:
            mb = rec.mfma_batches(grid, table_order=True)["batches"]
            for i in range(per):
                bt = mb[i % nb]
                ents += bt.t_count
                N.check(L.dips_hash_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(bt),
                                               ctypes.byref(eng.hg), ctypes.byref(eng.opt), inv_b, N.ptr(eng.params),
                                               N.ptr(eng.m), N.ptr(eng.v), N.ptr(eng.count), N.ptr(eng.loss_hist),
                                               int(eng.loss_hist.numel()), N.ptr(eng._work), int(eng._work.numel()),
                                               N.ptr(eng._dwork), int(eng._dwork.numel()), stream),
                        "dips_hash_train_step")
        return ents

    run_steps(W)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    ents = run_steps(K)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    value = K * B / (ms / 1e3)
    step_s = ms / 1e3 / K
    bytes_step = eng.P * 48.0 + 24.0 * ents / K
    achieved = bytes_step / step_s / 1e9
    result = {
        "metric": "hash network (hashgrid) training points/s",
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"sphere {n}^3 on [-1,1]^3, hash network (hashgrid 16 levels x 2 features, T 2^19, "
                               f"N_min 16, N_max 2048, CELU 16), batch {B}, 4 zoom levels, "
                               f"{eng.P} parameters",
                   "global_batch": B, "grid": n, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(f"hashgrid sphere{n}") if args.batch == 0 else None,
                     "kernel": "whole dips_hash_train_step (9 launches), compulsory bytes; traffic: the measured "
                               "HBM bytes of one step (its 9 kernels)",
                     "bytes_per_step": bytes_step, "entries_per_step": ents / K},
        "records_build_s": t_records,
        "loss_tail": eng.loss_hist[: int(eng.count.item())][-4:].cpu().numpy().tolist(),
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline("hashgrid", n, args.cpu_seconds, batch=256)
    print(json.dumps(result), flush=True)


def cpu_baseline(workload, n, seconds, batch=2048):
    """The reference's per-step pattern (oracle.refpattern, numpy float32) on
    the host cores: oracle.cpu_bench in a child interpreter, one
    single-threaded process per core (at most 16), each on its own stretch of
    batches through the middle of the grid, for ~``seconds``."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-m", "oracle.cpu_bench", "--workload", workload, "--n", str(n),
                        "--seconds", str(seconds), "--batch", str(batch)], cwd=ROOT, env=env, capture_output=True,
                       text=True,
                       timeout=4 * seconds + 300)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline failed:\n{r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


if __name__ == "__main__":
    main()
