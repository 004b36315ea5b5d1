"""Optimizer configuration (jax_dips/solvers/optimizers.py:13-97).

The update itself runs in the fused device kernel dips_opt_step; these
functions keep the reference's names and return the configuration it consumes.
"custom" = clip_by_global_norm(max_norm) -> scale_by_adam -> schedule -> scale(-1);
"adam" = optax.adam(lr); "rmsprop" = optax.rmsprop(lr).  Schedules:
exponential_decay / polynomial_schedule with transition_steps = 1000.
"""


def get_scheduler(scheduler_name="exponential", learning_rate=1e-2, decay_rate=0.96, transition_steps=1000, **kw):
    if scheduler_name not in ("exponential", "polynomial"):
        raise ValueError(f"Unknown scheduler: {scheduler_name}")
    return dict(scheduler_name=scheduler_name, learning_rate=learning_rate, decay_rate=decay_rate,
                transition_steps=transition_steps)


def chained_adam(scheduler_name="exponential", learning_rate=1e-2, decay_rate=0.96, transition_steps=1000,
                 max_norm=1.0, **kw):
    return dict(optimizer_name="custom", max_norm=max_norm,
                sched=get_scheduler(scheduler_name, learning_rate, decay_rate, transition_steps),
                learning_rate=learning_rate)


def get_optimizer(optimizer_name="custom", scheduler_name="exponential", learning_rate=1e-2, decay_rate=0.96,
                  max_norm=1.0, loss_fn=None, **kwargs):
    if optimizer_name == "custom":
        return chained_adam(scheduler_name=scheduler_name, learning_rate=learning_rate, decay_rate=decay_rate,
                            max_norm=max_norm)
    if optimizer_name in ("adam", "rmsprop"):
        return dict(optimizer_name=optimizer_name, learning_rate=learning_rate,
                    sched=dict(scheduler_name="exponential", decay_rate=decay_rate))
    raise ValueError("Unknown optimizer: {}".format(optimizer_name))
