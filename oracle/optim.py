"""optax 0.1.5 optimizer chains used by the trainer — CPU restatement.

Test infrastructure only (see oracle/__init__.py).

Restates jax_dips/solvers/optimizers.py:13-97:
  custom  = chain(clip_by_global_norm(1.0), scale_by_adam(0.9, 0.999, 1e-8),
                  scale_by_schedule(schedule), scale(-1.0))
  adam    = optax.adam(lr)          (scale_by_adam then -lr)
  rmsprop = optax.rmsprop(lr)       (decay 0.9, eps 1e-8, then -lr)
  schedule: exponential_decay(lr, transition_steps=1000, decay_rate) (smooth,
  count = optimizer steps taken so far) or polynomial_schedule(lr, 0, 1, 1000).

Pinned by tests/test_oracle_optim.py: scale_by_adam against torch.optim.Adam
(lr = 1, no weight decay) on the same gradients; clip and schedules against
their closed forms.
"""
import numpy as np


class OptState:
    def __init__(self, n, dt):
        self.mu = np.zeros(n, dtype=dt)
        self.nu = np.zeros(n, dtype=dt)
        self.count = 0


def schedule(name, lr, decay_rate, count, dt, transition_steps=1000):
    if name == "exponential":
        if count <= 0:
            return dt(lr)
        return dt(lr) * np.power(dt(decay_rate), dt(count) / dt(transition_steps))
    if name == "polynomial":
        c = min(max(count, 0), transition_steps)
        frac = dt(1) - dt(c) / dt(transition_steps)
        return dt(lr) * frac
    raise ValueError(f"unknown scheduler {name}")


def step(params, grads, state, cfg, dt=np.float32):
    """One optimizer update; returns new params (the state is updated in place).
    cfg: dict(optimizer_name, learning_rate, scheduler_name, decay_rate)."""
    g = np.asarray(grads, dtype=dt)
    p = np.asarray(params, dtype=dt)
    name = cfg.get("optimizer_name", "custom")
    lr = cfg.get("learning_rate", 1e-3)
    if name == "custom":
        gn = np.sqrt(np.sum(g * g, dtype=dt), dtype=dt)
        if not gn < dt(1.0):
            g = (g / gn) * dt(1.0)
    if name in ("custom", "adam"):
        state.mu = dt(1 - 0.9) * g + dt(0.9) * state.mu
        state.nu = dt(1 - 0.999) * (g * g) + dt(0.999) * state.nu
        t = state.count + 1
        mu_hat = state.mu / (dt(1) - np.power(dt(0.9), dt(t)))
        nu_hat = state.nu / (dt(1) - np.power(dt(0.999), dt(t)))
        u = mu_hat / (np.sqrt(nu_hat) + dt(1e-8))
        if name == "custom":
            u = schedule(cfg.get("scheduler_name", "exponential"), lr, cfg.get("decay_rate", 0.9),
                         state.count, dt) * u
            u = u * dt(-1)
        else:
            u = u * dt(-lr)
    elif name == "rmsprop":
        state.nu = dt(1 - 0.9) * (g * g) + dt(0.9) * state.nu
        u = g / np.sqrt(state.nu + dt(1e-8))
        u = u * dt(-lr)
    else:
        raise ValueError(f"unknown optimizer {name}")
    state.count += 1
    return (p + u).astype(dt)
