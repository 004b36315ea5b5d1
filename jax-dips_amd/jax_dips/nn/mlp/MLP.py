"""Two-phase tanh MLP (DoubleMLP) description and parameter handling.

Reference: jax_dips/nn/mlp/MLP.py:36-139.  u(x) = NN+(x) where phi(x) >= 0 and
NN-(x) otherwise (MLP.py:98); NN = [Linear(H), tanh] * L followed by
Linear(1).  Parameters keep the reference's haiku tree names
``double_mlp/~mlp_{p,m}_fn/linear[_k]`` with ``{"w": [in, out], "b": [out]}``.
Device kernels see one flat float32 vector: the + network then the - network,
each as W0, b0, ..., Wout, bout (row-major W).
"""
import numpy as np

SUPPORTED_ACTIVATIONS = ("jnp.tanh", "nn.tanh", "tanh")


def _tn(rng, n, std):
    """haiku TruncatedNormal: N(0,1) truncated to [-2, 2], scaled by std / 0.8796..."""
    w = rng.standard_normal(n)
    while True:
        bad = np.abs(w) > 2.0
        if not bad.any():
            break
        w[bad] = rng.standard_normal(int(bad.sum()))
    return w * (std / 0.87962566103423978)


class DoubleMLP:
    def __init__(self, name=None, model_type="mlp", mlp=None, resnet=None, **kwargs):
        if model_type != "mlp":
            raise NotImplementedError(f"model_type {model_type!r}: only 'mlp' has device kernels")
        mlp = dict(mlp or {"hidden_layers_m": 1, "hidden_dim_m": 1, "activation_m": "jnp.tanh",
                           "hidden_layers_p": 2, "hidden_dim_p": 10, "activation_p": "jnp.tanh"})
        for k in ("activation_m", "activation_p"):
            if mlp.get(k, "jnp.tanh") not in SUPPORTED_ACTIVATIONS:
                raise NotImplementedError(f"{k}={mlp.get(k)!r}: the device kernels implement tanh only")
        self.Lm, self.Hm = int(mlp["hidden_layers_m"]), int(mlp["hidden_dim_m"])
        self.Lp, self.Hp = int(mlp["hidden_layers_p"]), int(mlp["hidden_dim_p"])

    @property
    def dims(self):
        return (self.Lm, self.Hm, self.Lp, self.Hp)

    def shapes(self, tag):
        L, H = (self.Lp, self.Hp) if tag == "p" else (self.Lm, self.Hm)
        return [(3, H)] + [(H, H)] * (L - 1) + [(H, 1)]

    def names(self, tag):
        n = len(self.shapes(tag))
        return [f"double_mlp/~mlp_{tag}_fn/linear" + ("" if i == 0 else f"_{i}") for i in range(n)]

    def init(self, rng):
        """Parameter tree: hidden weights TruncatedNormal(0.1), output weights
        TruncatedNormal(1/sqrt(fan_in)), biases 0 (MLP.py:65, 70, 112-138).  ``rng``
        is an int seed or numpy Generator (haiku's PRNGKey stream is not reproduced)."""
        if not isinstance(rng, np.random.Generator):
            rng = np.random.default_rng(0 if rng is None else int(np.asarray(rng).ravel()[-1]))
        tree = {}
        for tag in ("p", "m"):
            shp = self.shapes(tag)
            for i, ((a, b), name) in enumerate(zip(shp, self.names(tag))):
                std = 0.1 if i < len(shp) - 1 else 1.0 / np.sqrt(a)
                tree[name] = {"w": _tn(rng, a * b, std).reshape(a, b).astype(np.float32),
                              "b": np.zeros(b, np.float32)}
        return tree

    def flatten(self, tree):
        parts = []
        for tag in ("p", "m"):
            for (a, b), name in zip(self.shapes(tag), self.names(tag)):
                w = np.asarray(tree[name]["w"], np.float32).reshape(a, b)
                parts += [w.reshape(-1), np.asarray(tree[name]["b"], np.float32).reshape(b)]
        return np.concatenate(parts).astype(np.float32)

    def unflatten(self, flat):
        flat = np.asarray(flat, np.float32)
        tree, off = {}, 0
        for tag in ("p", "m"):
            for (a, b), name in zip(self.shapes(tag), self.names(tag)):
                w = flat[off:off + a * b].reshape(a, b).copy()
                off += a * b
                bb = flat[off:off + b].copy()
                off += b
                tree[name] = {"w": w, "b": bb}
        return tree
