"""MI355X-native NBM Poisson trainer with the JAX-DIPS Python contracts.

The package keeps the reference's import paths for the training step:
``jax_dips.solvers.poisson.trainer.setup``, ``jax_dips.domain.mesh``,
``jax_dips.nn.configure.get_model``.  All per-point work runs in the HIP
library ``libdips_hip.so`` (built from ``csrc/``); there is no CPU compute
fallback: a missing or unloadable library raises on first use.
"""
__version__ = "0.1.0"
