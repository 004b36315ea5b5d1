"""Host-side preconditioner parameter container (no GPU): sizes agree with the
C-ABI's parameter count and the oracle's flat layout, flatten/unflatten round
trip, flax-layout tree, and unsupported configurations raise."""
import ctypes

import numpy as np
import pytest

from oracle import precond as PC


@pytest.mark.parametrize("widths", [(8, 4), (8,), (3, 5)])
def test_sizes_and_round_trip(widths):
    from jax_dips import _native
    from jax_dips.nn.preconditioner import Preconditioner

    p = Preconditioner(Ds=list(widths), scaling_coeff=1.0)
    assert p.size == PC.size(widths)
    net = _native.Net(1, 1, 2, 10, widths[0], widths[1] if len(widths) > 1 else 0, 1.0)
    assert _native.lib().dips_param_count(ctypes.byref(net)) == 167 + p.size
    tree = p.init(7)
    assert sorted(tree["params"]) == [f"Dense_{i}" for i in range(len(widths) + 1)]
    assert tree["params"]["Dense_0"]["kernel"].shape == (26, widths[0])
    flat = p.flatten(tree)
    assert flat.size == p.size and flat.dtype == np.float32
    assert np.array_equal(p.flatten(p.unflatten(flat)), flat)
    # same flat order as the oracle: layer kernels row-major, then biases
    for (W, b), (a, c) in zip(PC.unflatten(flat, widths), p.shapes()):
        assert W.shape == (a, c) and b.shape == (c,)
    lim = np.sqrt(6.0 / (26 + widths[0]))
    assert np.all(np.abs(tree["params"]["Dense_0"]["kernel"]) <= lim)
    assert np.all(tree["params"]["Dense_0"]["bias"] == 0.0)


def test_unsupported_configurations():
    from jax_dips.nn.preconditioner import Preconditioner

    with pytest.raises(NotImplementedError):
        Preconditioner(Ds=[8, 4], out_dim=2)
    with pytest.raises(NotImplementedError):
        Preconditioner(Ds=[8, 4, 2])
    with pytest.raises(ValueError):
        Preconditioner(Ds=[8, 4]).unflatten(np.zeros(10, np.float32))
