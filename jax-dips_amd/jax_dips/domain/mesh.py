"""GridState and mesh construction (jax_dips/domain/mesh.py:40-208).

Same fields and constructors as the reference: ``construct(3)`` returns
``(init_fn, point_fn)`` where ``init_fn(x, y, z)`` builds the 'ij' meshgrid
point list R (C order) and the six boundary point sets.
"""
from typing import Callable, Tuple

import numpy as np

from .._jaxmd_modules import dataclasses
from .._jaxmd_modules.util import Array


@dataclasses.dataclass
class GridState(object):
    x: Array
    y: Array
    z: Array
    R: Array
    dx: float
    dy: float
    dz: float
    R_xmin_boundary: Array
    R_xmax_boundary: Array
    R_ymin_boundary: Array
    R_ymax_boundary: Array
    R_zmin_boundary: Array
    R_zmax_boundary: Array

    def shape(self):
        return self.x.shape + self.y.shape + self.z.shape

    def xmin(self):
        return self.x.min()

    def xmax(self):
        return self.x.max()

    def ymin(self):
        return self.y.min()

    def ymax(self):
        return self.y.max()

    def zmin(self):
        return self.z.min()

    def zmax(self):
        return self.z.max()

    def base_level(self):
        return int(np.log2(len(self.R) ** (1.0 / 3.0)))


def construct(dimension: int) -> Tuple[Callable, Callable]:
    def init_fn_2d(x, y):
        x, y = np.asarray(x), np.asarray(y)
        dx, dy = x[1] - x[0], y[1] - y[0]
        X_, Y_ = np.meshgrid(x, y, indexing="ij")
        R = np.column_stack((X_.flatten(), Y_.flatten()))
        return GridState(x, y, None, R, dx, dy, None,
                         np.column_stack((X_[0, :].flatten(), Y_[0, :].flatten())),
                         np.column_stack((X_[-1, :].flatten(), Y_[-1, :].flatten())),
                         np.column_stack((X_[:, 0].flatten(), Y_[:, 0].flatten())),
                         np.column_stack((X_[:, -1].flatten(), Y_[:, -1].flatten())), None, None)

    def init_fn_3d(x, y, z):
        x, y, z = np.asarray(x), np.asarray(y), np.asarray(z)
        dx, dy, dz = x[1] - x[0], y[1] - y[0], z[1] - z[0]
        X_, Y_, Z_ = np.meshgrid(x, y, z, indexing="ij")
        R = np.column_stack((X_.flatten(), Y_.flatten(), Z_.flatten()))

        def face(sl):
            return np.column_stack((X_[sl].flatten(), Y_[sl].flatten(), Z_[sl].flatten()))

        return GridState(x, y, z, R, dx, dy, dz,
                         face((0, slice(None), slice(None))), face((-1, slice(None), slice(None))),
                         face((slice(None), 0, slice(None))), face((slice(None), -1, slice(None))),
                         face((slice(None), slice(None), 0)), face((slice(None), slice(None), -1)))

    def point3d_at(gstate, idx):
        return [gstate.x[idx[0]], gstate.y[idx[1]], gstate.z[idx[2]]]

    def point2d_at(gstate, idx):
        return [gstate.x[idx[0]], gstate.y[idx[1]]]

    if dimension == 3:
        return init_fn_3d, point3d_at
    if dimension == 2:
        return init_fn_2d, point2d_at
    raise ValueError(f"unsupported dimension {dimension}")


def init_gstate_3d_manually(x, y, z, dx, dy, dz, R_xmin_boundary, R_xmax_boundary, R_ymin_boundary,
                            R_ymax_boundary, R_zmin_boundary, R_zmax_boundary):
    R = np.column_stack((x, y, z))
    return GridState(x, y, z, R, dx, dy, dz, R_xmin_boundary, R_xmax_boundary, R_ymin_boundary,
                     R_ymax_boundary, R_zmin_boundary, R_zmax_boundary)
