"""The multiresolution hash network model type: two coordinate MLPs (one per
phase) on a position encoding, the identity (x itself) or the multiresolution
hashgrid.

Reference: make_hash_network / SolutionMLP (jax_dips/nn/hash_encoding/model.py:
53-161): u(x) = sol_act(sol_p(enc(x))) where phi(x) >= 0, sol_act(sol_m(enc(x)))
otherwise (model.py:89-94), each sol_* a CoordinateBasedMLP (blocks/nerfs_flax.py:
142-170): [Dense(D_i), celu] for the widths D_i, then Dense(1), glorot-uniform
kernels and zero biases.  ``sol_act`` "linear" (the default) is the identity.

* ``pos_enc="identity"`` (model.py:120-121, the setting of the reference's LPBE
  configuration, examples/benchmark_LPBE/conf/lpbe.yaml:48-60): the device
  kernels run it as a two-phase network with CELU hidden activations
  (dips_net_t.act = 1) and the flat layout of the tanh networks (the + network
  sol_p, then sol_m; per layer kernel [in][out] row-major, then bias); hidden
  widths [16] or [16, 16].
* ``pos_enc="hashgrid"`` (model.py:124-134, HashGridEncoder, blocks/
  encoders_flax.py:67-254, with N_max scaled by the bound): L = pos_levels
  levels of F = 2 features; the encoder's table is the parameter
  "latent codes stored on grid vertices" of the "position_encoder" module,
  initialised U(-1e-4, 1e-4).  Device kernels: dips_hash_* (one hidden layer
  of width 16).  Flat layout: sol_p, sol_m, then the table (rows x F).

Parameter tree (flax): {"params": {"sol_p": {"Dense_i": {"kernel", "bias"}},
"sol_m": {...}[, "position_encoder": {"latent codes stored on grid vertices":
(rows, F)}]}}.

* ``pos_enc="frequency"`` (model.py:122-123, FrequencyEncoder(L=10),
  blocks/encoders_flax.py:269-307): sin and cos of 2^l pi pos per axis and
  l < 10, 60 features; the same device kernels (dips_hashgrid_t.encoder = 1),
  no table, one hidden layer of 16.
* ``pos_enc="spherical_harmonics"`` (model.py:135-136,
  SphericalHarmonicsEncoder(L=highest_sh_order), blocks/encoders_flax.py:
  311-538): the real spherical harmonics of orders below highest_sh_order
  (1..8, highest_sh_order^2 features) on the raw position; the same device
  kernels (dips_hashgrid_t.encoder = 2), no table, one hidden layer of 16.

Not implemented (NotImplementedError): skip connections, output activations
other than linear, other hidden widths, and the total variation term (it does
not enter the solution, model.py:82-94).
"""
import numpy as np

SUPPORTED_WIDTHS = ([16], [16, 16])
HASHGRID_WIDTHS = ([16],)
FREQUENCIES = 10
# encoders that train through dips_hash_* (the identity encoder runs in the tile kernels)
DENSE_ENCODERS = ("hashgrid", "frequency", "spherical_harmonics")
TABLE_KEY = "latent codes stored on grid vertices"


class HashNetwork:
    act = 1  # CELU in the device kernels

    def __init__(self, bound=1.0, pos_enc="identity", tv_scale=0.0, pos_levels=16, layer_widths=(64,),
                 sol_skip_in_layers=(), sol_act="linear", highest_sh_order=8, N_min=16, N_max=2048, F=2,
                 T=2 ** 19, **kwargs):
        if pos_enc not in ("identity",) + DENSE_ENCODERS:
            raise NotImplementedError(f"pos_enc={pos_enc!r}: only the identity, hashgrid, frequency and "
                                      "spherical_harmonics encoders have device kernels")
        if list(sol_skip_in_layers):
            raise NotImplementedError("skip connections have no device implementation")
        if sol_act != "linear":
            raise NotImplementedError(f"sol_act={sol_act!r}: only 'linear' has a device implementation")
        widths = [int(w) for w in layer_widths]
        allowed = HASHGRID_WIDTHS if pos_enc in DENSE_ENCODERS else SUPPORTED_WIDTHS
        if widths not in allowed:
            raise NotImplementedError(f"layer_widths={widths}: compiled for {allowed} with pos_enc={pos_enc!r}")
        self.bound = float(bound)
        self.widths = widths
        self.L, self.H = len(widths), widths[0]
        self.encoder = pos_enc
        self.in_dim = 3
        self.table_rows = 0
        if pos_enc == "hashgrid":
            if int(F) != 2:
                raise NotImplementedError(f"F={F}: the hashgrid kernels use 2 features per table row")
            if not 2 <= int(pos_levels) <= 16:
                raise NotImplementedError(f"pos_levels={pos_levels}: the hashgrid kernels take 2..16 levels")
            # (levels, features, table_size, n_min, n_max, bound, encoder) of dips_hashgrid_t;
            # the encoder's finest resolution is int(N_max * bound) (model.py:129)
            self.hashgrid = (int(pos_levels), int(F), int(T), int(N_min), int(N_max * bound), self.bound, 0)
            self.in_dim = int(pos_levels) * int(F)
            self.table_rows = self._table_rows()
        elif pos_enc == "frequency":
            # FrequencyEncoder(L=10) (model.py:122-123): sin and cos of 2^l pi pos per axis
            self.hashgrid = (FREQUENCIES, 2, 1, 1, 1, self.bound, 1)
            self.in_dim = 6 * FREQUENCIES
        elif pos_enc == "spherical_harmonics":
            # SphericalHarmonicsEncoder(L=highest_sh_order) (model.py:135-136): orders below highest_sh_order
            order = int(highest_sh_order)
            if not 1 <= order <= 8:
                raise NotImplementedError(f"highest_sh_order={highest_sh_order}: the reference implements 1..8")
            self.hashgrid = (order, 2, 1, 1, 1, self.bound, 2)
            self.in_dim = order * order

    def _table_rows(self):
        """Rows of the table: per level min(res^3, T) (encoders_flax.py:98-112)."""
        import math

        L, _, T, nmin, nmax, _, _ = self.hashgrid
        b = math.exp((math.log(nmax) - math.log(nmin)) / (L - 1))
        rows = 0
        for i in range(L):
            res = math.ceil(nmin * b ** i - 1) + 1
            rows += min(res ** 3, T)
        return rows

    @property
    def dims(self):
        if self.encoder in DENSE_ENCODERS:
            raise AttributeError("the hashgrid network has no (Lm, Hm, Lp, Hp) shape; it trains through dips_hash_*")
        return (self.L, self.H, self.L, self.H)

    @property
    def n_params(self):
        return (sum(a * b + b for tag in ("p", "m") for a, b in self.shapes(tag))
                + self.table_rows * (self.hashgrid[1] if self.encoder == "hashgrid" else 0))

    def shapes(self, tag):
        dims = [self.in_dim] + self.widths + [1]
        return [(dims[i], dims[i + 1]) for i in range(len(dims) - 1)]

    def init(self, rng):
        """Glorot-uniform kernels, zero biases (nerfs_flax.py:151, flax Dense),
        table U(-1e-4, 1e-4) (encoders_flax.py:118-123); numpy's generator
        stands in for flax's PRNG stream."""
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
        if self.encoder == "hashgrid":
            F = self.hashgrid[1]
            params["position_encoder"] = {
                TABLE_KEY: rng.uniform(-1e-4, 1e-4, size=(self.table_rows, F)).astype(np.float32)}
        return {"params": params}

    def flatten(self, tree):
        t = tree["params"] if "params" in tree else tree
        parts = []
        for tag in ("p", "m"):
            for i, (a, b) in enumerate(self.shapes(tag)):
                d = t[f"sol_{tag}"][f"Dense_{i}"]
                parts += [np.asarray(d["kernel"], np.float32).reshape(a * b),
                          np.asarray(d["bias"], np.float32).reshape(b)]
        if self.encoder == "hashgrid":
            tab = np.asarray(t["position_encoder"][TABLE_KEY], np.float32)
            if tab.shape != (self.table_rows, self.hashgrid[1]):
                raise ValueError(f"table shape {tab.shape}, expected {(self.table_rows, self.hashgrid[1])}")
            parts.append(tab.reshape(-1))
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
        if self.encoder == "hashgrid":
            F = self.hashgrid[1]
            params["position_encoder"] = {TABLE_KEY: flat[off:off + self.table_rows * F].reshape(-1, F).copy()}
        return {"params": params}
