"""Network evaluation on point sets through dips_eval (u, grad u, d_n u).

Reference: Trainer.evaluate_solution_fn and evaluate_solution_and_gradients
(jax_dips/solvers/poisson/trainer.py:836-977).
"""
import ctypes

import numpy as np

from . import _native as N


def evaluate_points(model, flat_params, X, phase_p, normals=None, want_grad=False, device="cuda"):
    """u (m,) [and grad (m,3), grad.n (m,)] of the two-phase network at X (m,3)."""
    import torch

    if getattr(model, "encoder", "identity") in ("hashgrid", "frequency", "spherical_harmonics"):
        return _evaluate_hash(model, flat_params, X, phase_p, normals, want_grad, device)
    net = N.Net(*model.dims, 0, 0, 0.0, int(getattr(model, "act", 0)))
    X = torch.as_tensor(np.asarray(X, np.float32), device=device).contiguous()
    m = X.shape[0]
    ph = torch.as_tensor(np.asarray(phase_p, dtype=np.uint8), device=device).contiguous()
    prm = torch.as_tensor(np.asarray(flat_params, np.float32), device=device).contiguous()
    u = torch.empty(m, dtype=torch.float32, device=device)
    g = torch.empty((m, 3), dtype=torch.float32, device=device) if (want_grad or normals is not None) else None
    nr = None
    gn = None
    if normals is not None:
        nr = torch.as_tensor(np.asarray(normals, np.float32), device=device).contiguous()
        gn = torch.empty(m, dtype=torch.float32, device=device)
    N.check(N.lib().dips_eval(ctypes.byref(net), N.ptr(prm), N.ptr(X), N.ptr(ph), N.ptr(nr), m, N.ptr(u), N.ptr(g),
                              N.ptr(gn), N.current_stream()), "dips_eval")
    out_u = u.cpu().numpy()
    if normals is not None:
        return out_u, g.cpu().numpy(), gn.cpu().numpy()
    if want_grad:
        return out_u, g.cpu().numpy()
    return out_u, None


def _evaluate_hash(model, flat_params, X, phase_p, normals, want_grad, device):
    """evaluate_points for the hashgrid / frequency encoders (dips_hash_eval);
    the normal derivative is grad u . n."""
    import torch

    hg = N.HashGrid(*model.hashgrid)
    X = torch.as_tensor(np.asarray(X, np.float32), device=device).contiguous()
    m = X.shape[0]
    ph = torch.as_tensor(np.asarray(phase_p, dtype=np.uint8), device=device).contiguous()
    prm = torch.as_tensor(np.asarray(flat_params, np.float32), device=device).contiguous()
    u = torch.empty(m, dtype=torch.float32, device=device)
    g = torch.empty((m, 3), dtype=torch.float32, device=device) if (want_grad or normals is not None) else None
    N.check(N.lib().dips_hash_eval(ctypes.byref(hg), N.ptr(prm), N.ptr(X), N.ptr(ph), m, N.ptr(u), N.ptr(g),
                                   N.current_stream()), "dips_hash_eval")
    out_u = u.cpu().numpy()
    if normals is not None:
        nr = torch.as_tensor(np.asarray(normals, np.float32), device=device)
        return out_u, g.cpu().numpy(), (g * nr).sum(dim=1).cpu().numpy()
    if want_grad:
        return out_u, g.cpu().numpy()
    return out_u, None


def central_normals(phi_fn, R, dx, dy, dz):
    """normal_point_fn (discretization.py:199-218): central differences of phi at
    the given spacing, normalised (float32 positions, float64 samples)."""
    from .level_records import sample
    import torch

    P = torch.as_tensor(np.asarray(R, np.float32), device="cuda")
    h = torch.tensor([dx, dy, dz], dtype=torch.float32, device="cuda")
    comps = []
    for a in range(3):
        e = torch.zeros(3, dtype=torch.float32, device="cuda")
        e[a] = 1.0
        comps.append((sample(phi_fn, P + e * h[a]) - sample(phi_fn, P - e * h[a])) / (2.0 * h[a].double()))
    g = torch.stack(comps, dim=1)
    n = g / torch.sqrt((g * g).sum(dim=1, keepdim=True))
    return n.to(torch.float32).cpu().numpy()
