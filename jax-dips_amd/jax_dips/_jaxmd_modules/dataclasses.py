"""Frozen dataclass helper with ``replace`` (jax_dips/_jaxmd_modules/dataclasses.py:26-62)."""
import dataclasses as _dc

replace = _dc.replace


def dataclass(clz):
    """A frozen dataclass, as the reference's pytree dataclass (without pytree registration)."""
    return _dc.dataclass(frozen=True)(clz)
