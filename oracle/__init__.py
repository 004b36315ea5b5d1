"""CPU oracle for the NBM Poisson training step — TEST INFRASTRUCTURE ONLY.

This package is a plain numpy restatement of the JAX-DIPS reference algorithm for
the neural-bootstrapping Poisson training step.  It exists so that the HIP
implementation can be checked against an independent computation.  Only the
test suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker / the timed CPU baseline;
the product path never calls into it.

Modules and the reference files they restate (paths relative to the JAX-DIPS
repository root):

* ``geometry``       — jax_dips/geometry/geometric_integrations_per_point.py:32-1009
* ``discretization`` — jax_dips/solvers/poisson/discretization.py:164-519
                       (per-point residual as an affine map of network outputs)
* ``mlp``            — jax_dips/nn/mlp/MLP.py:93-139 (DoubleMLP forward + hand VJP)
* ``precond``        — jax_dips/nn/preconditioner.py:10-35 and discretization.py:339, :418-419
                       (learned per-point preconditioner P = 0.5 + s*sigmoid(MLP(26 coefficients)))
* ``optim``          — jax_dips/solvers/optimizers.py:13-54 (optax 0.1.5 chain)
* ``trainer``        — jax_dips/solvers/poisson/trainer.py:501-591, 783-789, 892-912
                       and jax_dips/data/data_management.py:70-160, 319-326
* ``judge``          — jax_dips/solvers/poisson/discretization.py:522-741 with
                       jax_dips/domain/interpolate.py:762-816 (grid-value operator)
* ``refpattern``     — the reference's per-step evaluation pattern (geometry and
                       197 network positions per point recomputed every step),
                       used only as the timed CPU baseline.

Pinning status of each function is recorded in the module docstrings; every
function here is checked by tests/test_oracle_*.py against closed forms,
invariants, brute force or the reference's own known-answer test.
"""
