"""Array namespace for user callbacks that runs on numpy or torch arrays.

User callbacks (phi, mu, k, f, alpha, beta, g) are written once against this
namespace, e.g. ``lambda r: xnp.sqrt(r[0]**2 + r[1]**2 + r[2]**2) - 0.5``, and
receive coordinate-first arrays ``r`` of shape (3, m).  Called with numpy
arrays they run on the host; called with torch tensors they run wherever the
tensor lives (the trainer samples them on the GPU when torch sees one).
This replaces ``jax.numpy`` inside callbacks, since jax is not part of the
stack.
"""
import math

import numpy as np

pi = math.pi


def _is_torch(x):
    t = type(x)
    return t.__module__.startswith("torch") and t.__name__ == "Tensor"


def _first(args):
    for a in args:
        if _is_torch(a):
            return True
    return False


def _mk(np_fn, torch_name):
    def fn(*args, **kw):
        if _first(args):
            import torch

            targs = []
            ref = next(a for a in args if _is_torch(a))
            for a in args:
                targs.append(a if _is_torch(a) else torch.as_tensor(a, dtype=ref.dtype, device=ref.device))
            return getattr(torch, torch_name)(*targs, **kw)
        return np_fn(*args, **kw)

    fn.__name__ = torch_name
    return fn


sqrt = _mk(np.sqrt, "sqrt")
exp = _mk(np.exp, "exp")
log = _mk(np.log, "log")
sin = _mk(np.sin, "sin")
cos = _mk(np.cos, "cos")
tan = _mk(np.tan, "tan")
tanh = _mk(np.tanh, "tanh")
sinh = _mk(np.sinh, "sinh")
cosh = _mk(np.cosh, "cosh")
arctan2 = _mk(np.arctan2, "atan2")
arccos = _mk(np.arccos, "acos")
arcsin = _mk(np.arcsin, "asin")
arctan = _mk(np.arctan, "atan")
abs = _mk(np.abs, "abs")
sign = _mk(np.sign, "sign")
minimum = _mk(np.minimum, "minimum")
maximum = _mk(np.maximum, "maximum")
floor = _mk(np.floor, "floor")
isfinite = _mk(np.isfinite, "isfinite")


def where(c, a, b):
    if _first((c, a, b)):
        import torch

        tens = [x for x in (a, b, c) if _is_torch(x)]
        flt = [x for x in tens if x.is_floating_point()]
        dtype = flt[0].dtype if flt else torch.float64
        dev = tens[0].device
        a = a if _is_torch(a) else torch.as_tensor(a, dtype=dtype, device=dev)
        b = b if _is_torch(b) else torch.as_tensor(b, dtype=dtype, device=dev)
        return torch.where(c, a, b)
    return np.where(c, a, b)


def nan_to_num(x, nan=0.0):
    if _is_torch(x):
        import torch

        return torch.nan_to_num(x, nan=nan)
    return np.nan_to_num(x, nan=nan)


def zeros_like(x):
    if _is_torch(x):
        import torch

        return torch.zeros_like(x)
    return np.zeros_like(x)


def full_like(x, v):
    if _is_torch(x):
        import torch

        return torch.full_like(x, v)
    return np.full_like(x, v)
