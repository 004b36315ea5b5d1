"""Grid-value solve of the same cut-cell discretisation: fallback and judge.

``GridJudge(gstate, lvl_gstate, callbacks)`` builds the zoom-0 records of the
eval grid on the GPU (spacing = grid spacing) and exposes the reference's
``residual_vector(u)`` (trainer.py:883-890) and a matrix-free BiCGSTAB solve
(dips_bicgstab) of residual_vector(u) = 0.  The accuracy of a trained network
is judged against this grid solution (and, where one exists, the analytic
solution).

Deviation from the reference, by design: the reference reads grid values
through its quadratic interpolant, whose index clamp at the lower box faces
(interpolate.py:608-610) turns index-0 lookups into an extrapolation from
interior nodes, which makes residual_vector non-affine there.  The judge reads
node values directly (with the reference's linear ghost layer beyond the box),
so it is affine everywhere and identical to the reference elsewhere.
"""
import ctypes

import numpy as np

from . import _native as N
from .level_records import build_level
from .records import DeviceGrid


class GridJudge:
    def __init__(self, gstate, lvl_gstate, callbacks, device="cuda"):
        import torch

        self.grid = DeviceGrid(gstate.x, gstate.y, gstate.z, device=device)
        self.shape = self.grid.shape
        self.n = int(np.prod(self.shape))
        h = np.array([gstate.dx, gstate.dy, gstate.dz], np.float32)
        lv = lvl_gstate
        bounds = (lv.xmin(), lv.xmax(), lv.ymin(), lv.ymax(), lv.zmin(), lv.zmax())
        self.rec = build_level(self.grid, h, bounds, callbacks, self.n)
        self.work = torch.empty(6 * self.n, dtype=torch.float32, device=device)
        self.dwork = torch.empty(520, dtype=torch.float64, device=device)
        self.device = device

    def residual_vector(self, u):
        """r = lhs/diag - rhs/diag at every node for the grid function u."""
        import torch

        u = torch.as_tensor(np.asarray(u, np.float32).reshape(-1), device=self.device).contiguous()
        y = torch.empty(self.n, dtype=torch.float32, device=self.device)
        N.check(N.lib().dips_grid_apply(ctypes.byref(self.grid.c), ctypes.byref(self.rec.c), N.ptr(u), N.ptr(y), 1,
                                        N.current_stream()), "dips_grid_apply")
        return y.cpu().numpy()

    def solve(self, x0=None, max_iter=5000, rtol=1e-6):
        """BiCGSTAB solution (grid vector), iterations, relative true residual."""
        import torch

        x = torch.zeros(self.n, dtype=torch.float32, device=self.device) if x0 is None else \
            torch.as_tensor(np.asarray(x0, np.float32).reshape(-1), device=self.device).clone()
        it = ctypes.c_int32(0)
        rr = ctypes.c_double(0.0)
        N.check(N.lib().dips_bicgstab(ctypes.byref(self.grid.c), ctypes.byref(self.rec.c), N.ptr(x), N.ptr(self.work),
                                      N.ptr(self.dwork), int(max_iter), float(rtol), ctypes.byref(it), ctypes.byref(rr),
                                      N.current_stream()), "dips_bicgstab")
        return x.cpu().numpy(), it.value, rr.value
