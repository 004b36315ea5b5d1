"""The multiresolution hash network model type with the identity position
encoder: two coordinate MLPs (one per phase) on x itself.

Reference: make_hash_network / SolutionMLP (jax_dips/nn/hash_encoding/model.py:
53-161) with ``pos_enc="identity"`` (model.py:139-140, the setting of the
reference's LPBE configuration, examples/benchmark_LPBE/conf/lpbe.yaml:48-60):
u(x) = sol_act(sol_p(x)) where phi(x) >= 0, sol_act(sol_m(x)) otherwise
(model.py:89-94), each sol_* a CoordinateBasedMLP (blocks/nerfs_flax.py:142-170):
[Dense(D_i), celu] for the widths D_i, then Dense(1), glorot-uniform kernels and
zero biases.  ``sol_act`` "linear" (the default) is the identity.

Device kernels run it as a two-phase network with CELU hidden activations
(dips_net_t.act = 1) and the flat layout of the tanh networks: the + network
(sol_p) then the - network (sol_m), each per layer kernel [in][out] row-major
then bias.  Parameter tree (flax): {"params": {"sol_p": {"Dense_i": {"kernel",
"bias"}}, "sol_m": {...}}}.

Not implemented (NotImplementedError): the hashgrid, frequency and spherical
harmonics encoders, skip connections, output activations other than linear,
and hidden widths other than one or two layers of 16.
"""
import numpy as np

SUPPORTED_WIDTHS = ([16], [16, 16])


class HashNetwork:
    act = 1  # CELU in the device kernels

    def __init__(self, bound=1.0, pos_enc="identity", tv_scale=0.0, pos_levels=16, layer_widths=(64,),
                 sol_skip_in_layers=(), sol_act="linear", highest_sh_order=8, N_min=16, N_max=2048, F=2,
                 T=2 ** 19, **kwargs):
        if pos_enc != "identity":
            raise NotImplementedError(f"pos_enc={pos_enc!r}: only the identity encoder has device kernels")
        if list(sol_skip_in_layers):
            raise NotImplementedError("skip connections have no device implementation")
        if sol_act != "linear":
            raise NotImplementedError(f"sol_act={sol_act!r}: only 'linear' has a device implementation")
        widths = [int(w) for w in layer_widths]
        if widths not in SUPPORTED_WIDTHS:
            raise NotImplementedError(f"layer_widths={widths}: compiled for {SUPPORTED_WIDTHS}")
        self.bound = float(bound)
        self.widths = widths
        self.L, self.H = len(widths), widths[0]

    @property
    def dims(self):
        return (self.L, self.H, self.L, self.H)

    def shapes(self, tag):
        dims = [3] + self.widths + [1]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    def init(self, rng):
        """Glorot-uniform kernels, zero biases (nerfs_flax.py:151, flax Dense);
        numpy's generator stands in for flax's PRNG stream."""
        if not isinstance(rng, np.random.Generator):
            rng = np.random.default_rng(0 if rng is None else int(np.asarray(rng).ravel()[-1]))
        params = {}
        for tag in ("p", "m"):
            layers = {}
            for i, (a, b) in enumerate(self.shapes(tag)):
                lim = np.sqrt(6.0 / (a + b))
                layers[f"Dense_{i}"] = {"kernel": rng.uniform(-lim, lim, size=(a, b)).astype(np.float32),
                                        "bias": np.zeros(b, np.float32)}
            params[f"sol_{tag}"] = layers
        return {"params": params}

    def flatten(self, tree):
        t = tree["params"] if "params" in tree else tree
        parts = []
        for tag in ("p", "m"):
            for i, (a, b) in enumerate(self.shapes(tag)):
                d = t[f"sol_{tag}"][f"Dense_{i}"]
                parts += [np.asarray(d["kernel"], np.float32).reshape(a * b),
                          np.asarray(d["bias"], np.float32).reshape(b)]
        return np.concatenate(parts).astype(np.float32)

    def unflatten(self, flat):
        flat = np.asarray(flat, np.float32)
        params, off = {}, 0
        for tag in ("p", "m"):
            layers = {}
            for i, (a, b) in enumerate(self.shapes(tag)):
                k = flat[off:off + a * b].reshape(a, b).copy()
                off += a * b
                layers[f"Dense_{i}"] = {"kernel": k, "bias": flat[off:off + b].copy()}
                off += b
            params[f"sol_{tag}"] = layers
        return {"params": params}
