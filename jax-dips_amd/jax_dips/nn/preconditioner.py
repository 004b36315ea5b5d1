"""Learned per-point preconditioner: parameter container.

P(c) = 0.5 + scaling_coeff * sigmoid(MLP(c)) on the 26 cell coefficients c of a
training point, MLP = [Dense(D_i), tanh]* then Dense(out_dim)
(jax_dips/nn/preconditioner.py:10-35).  The residual of the point is
multiplied by P; P and its gradient are evaluated inside the tile training
kernel, so this module only holds, initialises and (un)flattens the parameters.

Parameter tree (flax's layout, as trainer.py:234-241 stores it under
params["preconditioner"]): {"params": {"Dense_i": {"kernel": (fan_in, fan_out),
"bias": (fan_out,)}}}.  Flat order: per layer the kernel row-major, then the
bias; the flat block follows the two phase networks' parameters.
"""
import numpy as np

N_COEFFS = 26


class Preconditioner:
    def __init__(self, Ds, out_dim=1, scaling_coeff=1.0):
        self.Ds = [int(d) for d in Ds]
        if int(out_dim) != 1:
            raise NotImplementedError("the preconditioner multiplies a scalar residual: out_dim must be 1")
        if not 1 <= len(self.Ds) <= 2 or min(self.Ds) < 1:
            raise NotImplementedError(f"preconditioner layer widths {self.Ds}: one or two hidden layers supported")
        self.out_dim = 1
        self.scaling_coeff = float(scaling_coeff)

    def shapes(self):
        dims = [N_COEFFS] + self.Ds + [self.out_dim]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    @property
    def size(self):
        return sum(a * b + b for a, b in self.shapes())

    def init(self, seed=42):
        """Glorot-uniform kernels and zero biases (the flax defaults of
        preconditioner.py:22); numpy's generator, not jax's PRNG stream."""
        rng = np.random.default_rng(seed)
        layers = {}
        for i, (a, b) in enumerate(self.shapes()):
            lim = np.sqrt(6.0 / (a + b))
            layers[f"Dense_{i}"] = {"kernel": rng.uniform(-lim, lim, size=(a, b)).astype(np.float32),
                                    "bias": np.zeros(b, np.float32)}
        return {"params": layers}

    def flatten(self, tree):
        layers = tree["params"] if "params" in tree else tree
        parts = []
        for i, (a, b) in enumerate(self.shapes()):
            d = layers[f"Dense_{i}"]
            parts += [np.asarray(d["kernel"], np.float32).reshape(a * b), np.asarray(d["bias"], np.float32).reshape(b)]
        return np.concatenate(parts).astype(np.float32)

    def unflatten(self, flat):
        flat = np.asarray(flat, np.float32)
        if flat.size != self.size:
            raise ValueError(f"expected {self.size} preconditioner parameters, got {flat.size}")
        layers, off = {}, 0
        for i, (a, b) in enumerate(self.shapes()):
            k = flat[off:off + a * b].reshape(a, b).copy()
            off += a * b
            layers[f"Dense_{i}"] = {"kernel": k, "bias": flat[off:off + b].copy()}
            off += b
        return {"params": layers}

    def engine_config(self):
        return {"widths": list(self.Ds), "scale": self.scaling_coeff}
