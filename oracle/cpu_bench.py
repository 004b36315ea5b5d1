"""CPU baseline timing of the reference's per-step pattern on all host cores.

Test infrastructure only (see oracle/__init__.py): bench.py starts this as a
child interpreter for its ``cpu_baseline`` figure.  Each of ``workers``
processes (single-threaded numpy) runs oracle.refpattern.step on consecutive
batches of ``batch`` points through the middle of the n^3 grid (level-0
spacing, its own disjoint stretch of the grid) for ``seconds``; the result is
the summed points per second, printed as one JSON line.

    python -m oracle.cpu_bench --workload sphere --n 128 --seconds 15 --workers 16
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for _p in (ROOT, os.path.join(ROOT, "jax-dips_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)


def _problem(workload):
    from jax_dips import problems

    if workload == "hashgrid":
        return problems.sphere(radius=0.5, mu_m=1.0, mu_p=10.0), None
    if workload == "pb":
        return problems.pb_box(n_atoms=2000, half_width=20.0, seed=0), (3, 64, 3, 64)
    if workload == "lpbe":
        return problems.lpbe_sphere(h=5.0 / 127), (3, 64, 3, 64)
    if workload == "star":
        dims = (3, 64, 3, 64)
        return problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3, mu_m=1.0, mu_p=100.0, k_m=1.0, k_p=1.0), dims
    return problems.sphere(radius=0.5, mu_m=1.0, mu_p=10.0), (1, 1, 2, 10)


def _worker(args):
    workload, n, seconds, batch, index, workers = args
    from jax_dips import problems
    from jax_dips.nn.mlp.MLP import DoubleMLP
    from oracle import discretization as D
    from oracle import refpattern

    prob, dims = _problem(workload)
    net = None
    if workload == "hashgrid":
        # the hash network at the reference's defaults (model.py:96-117)
        from jax_dips.nn.hash_network import HashNetwork
        from oracle import hashgrid as HG

        model = HashNetwork(bound=1.0, pos_enc="hashgrid", layer_widths=[16], pos_levels=16, T=2 ** 19, N_min=16,
                            N_max=2048, F=2)
        net = HG.Network(HG.HashGrid(16, 2 ** 19, 2, 16, 2048, 1.0), [16])
    else:
        mlp = {"hidden_layers_m": dims[0], "hidden_dim_m": dims[1], "activation_m": "jnp.tanh",
               "hidden_layers_p": dims[2], "hidden_dim_p": dims[3], "activation_p": "jnp.tanh"}
        model = DoubleMLP(mlp=mlp)
    theta = model.flatten(model.init(42))
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float32)
    half = {"pb": 20.0, "lpbe": 2.5}.get(workload, 1.0)
    x, y, z = problems.grid_axes(n, -half, half)
    h0 = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]], np.float32)
    bounds = (x.min(), x.max(), y.min(), y.max(), z.min(), z.max())
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    # worker k starts k stretches after the middle of the box (batches there cross the interface)
    start = P.shape[0] // 2 + index * (P.shape[0] // (4 * workers))
    done, t0 = 0, time.time()
    while time.time() - t0 < seconds:
        s = (start + done) % (P.shape[0] - batch)
        refpattern.step(theta, dims, P[s:s + batch], h0, bounds, cb, net=net)
        done += batch
    return done, time.time() - t0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="sphere", choices=["sphere", "star", "pb", "lpbe", "hashgrid"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--seconds", type=float, default=15.0)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--workers", type=int, default=0, help="0: min(16, host cores available)")
    a = ap.parse_args()
    workers = a.workers or min(16, len(os.sched_getaffinity(0)))
    t0 = time.time()
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.map(_worker, [(a.workload, a.n, a.seconds, a.batch, k, workers) for k in range(workers)])
    wall = time.time() - t0
    done = sum(d for d, _ in res)
    per = max(t for _, t in res)
    print(json.dumps({"value": done / per, "unit": "points/s", "cores": workers, "kind": "port",
                      "sample": f"{done} points in batches of {a.batch} through the middle of the {a.n}^3 grid, "
                                f"level-0 spacing, {workers} processes x {per:.1f} s (wall {wall:.1f} s), numpy "
                                f"float32 restatement of the reference's per-step pattern"
                                + (" with the hashgrid network" if a.workload == "hashgrid" else "")
                                + ", one thread each"}),
          flush=True)


if __name__ == "__main__":
    main()
