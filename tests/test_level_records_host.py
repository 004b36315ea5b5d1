"""The record build's deduplicated phi sampling (level_records.stencil_corner_phi)
returns exactly the values of sampling phi at every stencil point and cell
corner directly, in the reference's float32 order of operations
(s = fl32(p + e h), corner = fl32(s + c h/2)), for grid nodes and for extra
(padding) points, at every zoom spacing; and it asks the callback for fewer
positions.  Host tensors, no GPU."""
import numpy as np
import pytest

from jax_dips import problems


@pytest.mark.parametrize("level", [0, 1, 2, 3])
def test_dedup_phi_equals_direct_sampling(level):
    import torch

    from jax_dips.level_records import CORNER_SIGNS, STENCIL, sample, stencil_corner_phi
    from jax_dips.records import DeviceGrid

    rng = np.random.default_rng(level)
    n = 11
    x, y, z = problems.grid_axes(n)
    grid = DeviceGrid(x, y[:9], z[:10], extra=rng.uniform(-1, 1, (7, 3)), device="cpu")
    h0 = np.array([x[1] - x[0], y[1] - y[0], z[1] - z[0]], np.float32)
    h = (h0 * np.float32(0.5) ** level).astype(np.float32)
    ht = torch.tensor(h)
    half = ht * 0.5
    prob = problems.star(r0=0.5, amp=0.1, lobes_t=5, lobes_p=3)
    calls = []

    def phi(r):
        calls.append(r.shape[1])
        return prob["lvl_set_fn"](r)

    P = grid.points()
    S = P[:, None, :] + torch.tensor(STENCIL, dtype=torch.float32)[None] * ht
    phis, phic = stencil_corner_phi(grid, S, ht, half, phi)
    asked = sum(calls)
    C = S[:, :, None, :] + torch.tensor(CORNER_SIGNS, dtype=torch.float32)[None, None] * half
    want_s = sample(prob["lvl_set_fn"], S)
    want_c = sample(prob["lvl_set_fn"], C)
    assert phis.shape == (grid.npts, 7) and phic.shape == (grid.npts, 7, 8)
    assert torch.equal(phis, want_s)
    assert torch.equal(phic, want_c)
    direct = grid.npts * 63
    print(f"level {level}: {asked} positions asked instead of {direct}")
    assert asked < direct
