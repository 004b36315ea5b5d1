"""Batching and the resolution schedule (jax_dips/data/data_management.py:70-326).

Training points are the level-set-independent grid points ``tr_gstate.R``;
batches are contiguous slices of B points (the device works on index ranges,
so the (nb, B, 3) copy of the reference is never materialised).  A short last
batch is padded, as in the reference, with random points rescaled to the data
range (generate_random_points_like); those points enter the loss.
"""
import numpy as np


def generate_random_points_like(x_data, num_missing_points, seed=0):
    """Normal draws rescaled to [min, max] of the data (data_management.py:70-76).
    numpy's generator replaces jax's PRNGKey(0), so the draws differ."""
    rng = np.random.default_rng(seed)
    dR = rng.standard_normal((num_missing_points, 3))
    lo, hi = np.min(x_data), np.max(x_data)
    return (lo + (hi - lo) * (dR - dR.min()) / (dR.max() - dR.min())).astype(np.float32)


class DatasetDict:
    """Contiguous batches of x_data (num_gpus = 1)."""

    def __init__(self, x_data, batch_size, num_gpus=1):
        self.x_data = np.asarray(x_data)
        self._len = len(self.x_data)
        self.batch_size = min(int(batch_size), self._len)
        self.num_gpus = num_gpus
        self.num_batches = -(-self._len // self.batch_size)
        self.last_batch_size = self._len % self.batch_size

    def padding(self):
        if self.last_batch_size == 0:
            return np.zeros((0, 3), np.float32)
        return generate_random_points_like(self.x_data, self.batch_size - self.last_batch_size)

    def get_batched_data(self):
        pad = self.padding()
        full = np.concatenate([self.x_data.astype(np.float32), pad])
        return full.reshape(self.num_batches, self.batch_size, 3)


class TrainData:
    """Training points and the zoom schedule of the stencil spacing."""

    def __init__(self, gstate, lvl_set_fn=None, refine=False, refine_lod=False, refine_normals=False,
                 v_cycle_period=4, rest_at_level=10):
        if refine or refine_lod or refine_normals:
            raise NotImplementedError("training-point refinement is not part of the device trainer")
        self.gstate = gstate
        self.train_points = gstate.R
        self.lvl_set_fn = lvl_set_fn
        self.alt_res = False
        self.base_level = gstate.base_level()

    def alternate_res_sequentially(self, num_epochs, epoch, train_dx, train_dy, train_dz):
        """h = h_tr * 0.5 ** (epoch // (num_epochs // 4)) (data_management.py:319-326);
        num_epochs < 4 divides by zero as in the reference."""
        zoom_lvl = epoch // (num_epochs // 4)
        s = np.float32(0.5) ** np.float32(zoom_lvl)
        return (np.float32(self.gstate.dx) * s, np.float32(self.gstate.dy) * s, np.float32(self.gstate.dz) * s)
