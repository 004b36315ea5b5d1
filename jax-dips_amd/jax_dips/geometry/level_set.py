"""Level-set helpers on the training path (jax_dips/geometry/level_set.py:34-48)."""
from .. import xnp


def perturb_level_set_fn(phi_fn):
    """phi + sign+-(phi) * 1e-10 with sign+-(0) = -1, so phi is never 0 and
    phi = 0 maps to Omega-."""
    EPS = 1.0e-10

    def perturbed_phi_fn(x):
        lvl = phi_fn(x)
        return lvl + xnp.where(lvl > 0, 1.0, -1.0) * EPS

    return perturbed_phi_fn
