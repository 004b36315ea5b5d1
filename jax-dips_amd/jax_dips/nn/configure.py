"""Model factory (jax_dips/nn/configure.py:28-50): "mlp" (DoubleMLP, tanh) and
"multiresolution_hash_network" with the identity, hashgrid, frequency or spherical-harmonics
encoder (CELU coordinate MLPs, nn/hash_network.py).

``get_model(model_dict, model_type)`` returns ``(forward, framework)``;
``forward.init(rng, x, phi_x)`` builds the parameter tree and
``forward.apply(params, rng, x, phi_x)`` evaluates the two-phase network on the
GPU (dips_eval) at one point (x: (3,)) or many (x: (m,3)).
"""
import numpy as np

from .hash_network import HashNetwork
from .mlp.MLP import DoubleMLP


class _Forward:
    def __init__(self, model_dict, model_type="mlp"):
        if model_type == "multiresolution_hash_network":
            self.model = HashNetwork(**dict(model_dict[model_type]))
        else:
            md = {k: v for k, v in dict(model_dict).items() if k in ("name", "model_type", "mlp", "resnet")}
            self.model = DoubleMLP(**md)

    def init(self, rng, x=None, phi_x=None):
        return self.model.init(rng)

    def apply(self, params, *args):
        """apply(params, rng, x, phi_x) (haiku) or apply(params, x, phi_x) (flax)."""
        from ..evaluation import evaluate_points

        x, phi_x = args[-2], args[-1]

        x = np.asarray(x, np.float32)
        single = x.ndim == 1
        X = x.reshape(-1, 3)
        phase = np.asarray(phi_x).reshape(-1) >= 0
        u, _ = evaluate_points(self.model, self.model.flatten(params), X, phase)
        return u.reshape(1) if single else u.reshape(-1, 1)


def get_model(model_dict, model_type: str = "mlp"):
    if model_type in ("mlp",):
        return _Forward(model_dict), "haiku"
    if model_type == "multiresolution_hash_network":
        return _Forward(model_dict, model_type), "flax"
    raise NotImplementedError(f"model_type {model_type!r} has no device implementation")
