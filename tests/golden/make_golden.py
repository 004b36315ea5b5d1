"""Writes the small oracle fixtures in this folder (run from the repo root:
python tests/golden/make_golden.py).

sphere8_records.npz: affine records of the 8^3 sphere problem (radius 0.5,
mu- = 1, mu+ = 10) at zoom levels 0 and 2, float64, as sparse (row, key, w)
triplets plus constants, and the loss / gradient of a fixed parameter vector.
The fixtures guard the oracle against unintended drift; the oracle itself is
pinned by the closed-form and known-answer tests in tests/test_oracle_*.py.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "jax-dips_amd"), os.path.join(ROOT, "tests")]

from helpers import grid_points  # noqa: E402
from jax_dips import problems  # noqa: E402
from oracle import discretization as D  # noqa: E402
from oracle import mlp as M  # noqa: E402
from oracle import trainer as T  # noqa: E402

DIMS = (1, 1, 2, 10)


def build():
    prob = problems.sphere()
    cb = D.Callbacks(problems.oracle_callbacks(prob), np.float64)
    axes, P, h, bounds = grid_points(8)
    theta = M.init_params(DIMS, seed=123)
    out = {"theta": theta}
    for lvl in (0, 2):
        rec = T.LevelRecords(P, T.level_spacing(h, lvl), bounds, cb)
        r, k = np.nonzero(rec.W)
        loss, grad, _ = T.loss_and_grad(theta.astype(np.float64), DIMS, rec, np.arange(P.shape[0]))
        out.update({f"L{lvl}_row": r.astype(np.int32), f"L{lvl}_key": k.astype(np.int16),
                    f"L{lvl}_w": rec.W[r, k], f"L{lvl}_c": rec.c, f"L{lvl}_loss": np.array(loss),
                    f"L{lvl}_grad": grad})
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "sphere8_records.npz"), **build())
    print("wrote", os.path.join(HERE, "sphere8_records.npz"))
